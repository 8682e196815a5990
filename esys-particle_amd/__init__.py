"""MI355X-native per-timestep DEM particle engine with the ESyS-Particle API.

The hot path (neighbour-table rebuild, interaction storage update, force laws,
integration) is hand-written HIP in ``csrc/`` compiled into ``libesp.so``; this
package exposes it through the reference's scripting interface (``lsm.py``).
"""
from .lsm import (  # noqa: F401
    BoundingBox, BrittleBeamPrms, BrittleBeamSCPrms, DampingPrms, GravityPrms, LinDampingPrms,
    LocalDampingPrms, LsmMpi, NRotElasticPrms, RotDampingPrms, RotLocalDampingPrms,
    NRotElasticWallPrms, NRotSphere, RotBondPrms, RotElasticPrms, RotFrictionPrms,
    RestartCheckPointPrms, RotSphere, Runnable, TaggedIdConnection, Vec3,
    InteractionScalarFieldSaverPrms, WallVectorFieldSaverPrms, NRotBondedWallPrms,
)
from ._native import Engine, load_hip_library, LIB_PATH  # noqa: F401
