"""Python facade mirroring the ESyS-Particle scripting API for one subdomain.

A script written against ``esys.lsm`` (``LsmMpi``, ``NRotElasticPrms``,
``RotFrictionPrms``, ``RotBondPrms``, ``BrittleBeamPrms``, ``GravityPrms``,
``NRotElasticWallPrms``, ``createInteractionGroup``, ``runTimeStep`` ...) drives
the MI355X engine through the C-ABI in ``include/esp.h``.  Method names, argument
names and defaults follow the reference's Boost.Python bindings
(Python/esys/lsm/LsmMpiPy.cpp:2140-3380, InteractionParamsPy.cpp:1216-2422).

The whole particle set of the model lives on one GPU: ``numWorkerProcesses`` and
``mpiDimList`` are accepted for source compatibility and ignored.
"""
import math

import numpy as np

from . import _native


# --------------------------------------------------------------------------
# value types (esys.lsm.util)
# --------------------------------------------------------------------------
class Vec3:
    """3-vector (Foundation/vec3.h)."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        if hasattr(x, "__len__") and not isinstance(x, str):
            x, y, z = x
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __getitem__(self, i):
        return (self.x, self.y, self.z)[i]

    def __len__(self):
        return 3

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def X(self):
        return self.x

    def Y(self):
        return self.y

    def Z(self):
        return self.z

    def __add__(self, o):
        return Vec3(self.x + o[0], self.y + o[1], self.z + o[2])

    def __sub__(self, o):
        return Vec3(self.x - o[0], self.y - o[1], self.z - o[2])

    def __mul__(self, s):
        return Vec3(self.x * s, self.y * s, self.z * s)

    __rmul__ = __mul__

    def norm(self):
        return math.sqrt(self.x * self.x + self.y * self.y + self.z * self.z)

    def __repr__(self):
        return "Vec3(%r, %r, %r)" % (self.x, self.y, self.z)


class BoundingBox:
    def __init__(self, minPt, maxPt):
        self.minPt = Vec3(minPt)
        self.maxPt = Vec3(maxPt)

    def getMinPt(self):
        return self.minPt

    def getMaxPt(self):
        return self.maxPt


class _Sphere:
    def __init__(self, id, posn, radius, mass=None):
        self.id = int(id)
        self.posn = Vec3(posn)
        self.radius = float(radius)
        self.mass = None if mass is None else float(mass)
        self.tag = 0

    def getId(self):
        return self.id

    def getTag(self):
        return self.tag

    def setTag(self, tag):
        self.tag = int(tag)

    def getPosn(self):
        return self.posn

    def getRadius(self):
        return self.radius

    def getMass(self):
        return self.mass


class NRotSphere(_Sphere):
    """NRotSphere(id, posn, radius, mass) (ParticlePy.cpp)."""


class RotSphere(_Sphere):
    """RotSphere(id, posn, radius, mass) (RotParticlePy.cpp:36-39)."""


class TaggedIdConnection:
    """Bond connection (id1, id2, tag) for createConnections."""

    def __init__(self, id1, id2, tag=0):
        self.id1, self.id2, self.tag = int(id1), int(id2), int(tag)

    def getP1Id(self):
        return self.id1

    def getP2Id(self):
        return self.id2

    def getTag(self):
        return self.tag


class Runnable:
    """Base class of pre/post time-step runnables (Foundation/Runnable.h)."""

    def run(self):
        raise NotImplementedError


# --------------------------------------------------------------------------
# interaction parameter classes (InteractionParamsPy.cpp)
# --------------------------------------------------------------------------
class InteractionPrms:
    def getName(self):
        return self.name


class NRotElasticPrms(InteractionPrms):
    """NRotElasticPrms(name, normalK, scaling=True) (InteractionParamsPy.cpp:195-226)."""

    def __init__(self, name, normalK, scaling=True):
        self.name, self.normalK, self.scaling = name, float(normalK), bool(scaling)


class RotElasticPrms(InteractionPrms):
    """RotElasticPrms(name, normalK, scaling=True) (RotElasticInteraction.cpp:23-30)."""

    def __init__(self, name, normalK, scaling=True):
        self.name, self.normalK, self.scaling = name, float(normalK), bool(scaling)


class RotFrictionPrms(InteractionPrms):
    """RotFrictionPrms(name, normalK, dynamicMu, staticMu, shearK, scaling=True,
    rigid=False, meanR_scaling=True) (InteractionParamsPy.cpp:772-833)."""

    def __init__(self, name, normalK, dynamicMu, staticMu, shearK, scaling=True,
                 rigid=False, meanR_scaling=True):
        self.name = name
        self.normalK, self.dynamicMu, self.staticMu, self.shearK = (
            float(normalK), float(dynamicMu), float(staticMu), float(shearK))
        self.scaling, self.rigid, self.meanR_scaling = bool(scaling), bool(rigid), bool(meanR_scaling)


class RotBondPrms(InteractionPrms):
    """RotBondPrms(name, normalK, shearK, torsionK, bendingK, normalBrkForce,
    shearBrkForce, torsionBrkForce, bendingBrkForce, tag, scaling=True,
    meanR_scaling=True, truncated=1.0) (InteractionParamsPy.cpp:414-548).

    ``perBondStrengths`` (extension): an (n, 4) array of (normal, shear, torsion,
    bending) break values, one row per connection of ``tag`` in creation order,
    which the reference can only express through its restart files."""

    def __init__(self, name, normalK, shearK, torsionK, bendingK, normalBrkForce,
                 shearBrkForce, torsionBrkForce, bendingBrkForce, tag, scaling=True,
                 meanR_scaling=True, truncated=1.0, perBondStrengths=None):
        self.name = name
        self.kr, self.ks, self.kt, self.kb = (float(normalK), float(shearK), float(torsionK),
                                              float(bendingK))
        self.max_n, self.max_s, self.max_t, self.max_b = (
            float(normalBrkForce), float(shearBrkForce), float(torsionBrkForce),
            float(bendingBrkForce))
        self.tag = int(tag)
        self.scaling, self.meanR_scaling = bool(scaling), bool(meanR_scaling)
        self.truncated = float(truncated)
        self.beta1 = 1.0
        self.beta2 = 1.0
        self.perBondStrengths = perBondStrengths

    def _c(self):
        return {"kr": self.kr, "ks": self.ks, "kt": self.kt, "kb": self.kb,
                "max_n": self.max_n, "max_s": self.max_s, "max_t": self.max_t,
                "max_b": self.max_b, "tag": self.tag, "scaling": int(self.scaling),
                "meanR_scaling": int(self.meanR_scaling), "truncated": self.truncated,
                "beta1": self.beta1, "beta2": self.beta2}


class BrittleBeamPrms(RotBondPrms):
    """BrittleBeamPrms(name, youngsModulus, poissonsRatio, cohesion, tanAngle, tag,
    meanR_scaling=True, truncated=1.0, beta1=1.0, beta2=1.0): stiffnesses and
    strengths from elastic moduli (ARotBondedInteraction.cpp:91-109,
    RotBondedInteraction.cpp:73-94).  As in the reference's worker
    (RotSubLattice.hpp:123-166) beta1/beta2 do not reach the bonds and stay 1."""

    def __init__(self, name, youngsModulus, poissonsRatio, cohesion, tanAngle, tag,
                 meanR_scaling=True, truncated=1.0, beta1=1.0, beta2=1.0, perBondStrengths=None):
        shear_modulus = youngsModulus / (2.0 * (1. + poissonsRatio))
        kr = math.pi * youngsModulus
        ks = math.pi * shear_modulus
        kb = math.pi * youngsModulus
        kt = math.pi * shear_modulus
        RotBondPrms.__init__(self, name, kr, ks, kt, kb, math.pi * cohesion / tanAngle,
                             math.pi * cohesion, math.pi * cohesion,
                             math.pi * cohesion / tanAngle, tag, True, meanR_scaling,
                             truncated, perBondStrengths)


class BrittleBeamSCPrms(InteractionPrms):
    """BrittleBeamSCPrms(name, youngsModulus, poissonsRatio, cohesion, frictionAngle,
    tensileCutoff, tag) (InteractionParamsPy.cpp:680-693, 1689-1730): elastic
    beam bonds (ARotBondedInteraction) that break by a truncated Mohr-Coulomb
    test on the mean stress of their two particles (BrittleBeamSC.cpp:103-148).
    frictionAngle in degrees.  Stiffnesses as ARotBondedIGP(E, nu)
    (ARotBondedInteraction.cpp:94-110)."""

    def __init__(self, name, youngsModulus, poissonsRatio, cohesion, frictionAngle,
                 tensileCutoff, tag):
        shear_modulus = youngsModulus / (2.0 * (1. + poissonsRatio))
        self.name = name
        self.youngsModulus, self.poissonsRatio = youngsModulus, poissonsRatio
        self.cohesion, self.frictionAngle, self.tensileCutoff = cohesion, frictionAngle, tensileCutoff
        self.tag = tag
        self.kr = math.pi * youngsModulus
        self.ks = math.pi * shear_modulus
        self.kb = math.pi * youngsModulus
        self.kt = math.pi * shear_modulus

    def _c(self):
        return {"kr": self.kr, "ks": self.ks, "kt": self.kt, "kb": self.kb,
                "cohesion": self.cohesion, "friction_angle": self.frictionAngle,
                "tension_cutoff": self.tensileCutoff, "tag": self.tag}


class RestartCheckPointPrms:
    """RestartCheckPointPrms(fileNamePrefix, beginTimeStep, endTimeStep,
    timeStepIncr, precision=12) (CheckPointParamsPy.cpp:140-160): restartable
    checkpoints after every step t with begin <= t <= end and
    (t - begin) % incr == 0, and after step end (CheckPointController.cpp:349-363)."""

    def __init__(self, fileNamePrefix, beginTimeStep, endTimeStep, timeStepIncr, precision=12):
        self.fileNamePrefix = fileNamePrefix
        self.beginTimeStep, self.endTimeStep = int(beginTimeStep), int(endTimeStep)
        self.timeStepIncr, self.precision = int(timeStepIncr), int(precision)

    def isCheckPoint(self, t):
        b, e, i = self.beginTimeStep, self.endTimeStep, self.timeStepIncr
        return (b <= t <= e and (t - b) % i == 0) or t == e

    def nextCheckPoint(self, t):
        """First checkpoint step after step t, or None."""
        b, e, i = self.beginTimeStep, self.endTimeStep, self.timeStepIncr
        if t < b:
            return b if b <= e else e
        if t >= e:
            return None
        return min(b + ((t - b) // i + 1) * i, e)


class FieldSaverPrms:
    """Common part of the field saver parameters (FieldSaverPrmsPy.cpp:46-70):
    a field is saved after every step t with begin <= t <= end and
    (t - begin) % incr == 0 (AFieldMaster::needSave, FieldMaster.cpp:97-100);
    step 0 is the state before the first step."""

    def __init__(self, fieldName, fileName, fileFormat, beginTimeStep, endTimeStep, timeStepIncr):
        self.fieldName, self.fileName, self.fileFormat = fieldName, fileName, fileFormat
        self.beginTimeStep, self.endTimeStep = int(beginTimeStep), int(endTimeStep)
        self.timeStepIncr = int(timeStepIncr)
        if self.timeStepIncr <= 0:
            raise ValueError("timeStepIncr must be positive")

    def needSave(self, t):
        b, e, i = self.beginTimeStep, self.endTimeStep, self.timeStepIncr
        return (t - b) % i == 0 and b <= t <= e

    def nextSave(self, t):
        """First save step after step t, or None."""
        b, e, i = self.beginTimeStep, self.endTimeStep, self.timeStepIncr
        nxt = b if t < b else b + ((t - b) // i + 1) * i
        return nxt if nxt <= e else None


class WallVectorFieldSaverPrms(FieldSaverPrms):
    """WallVectorFieldSaverPrms(wallName, fieldName, fileName, fileFormat,
    beginTimeStep, endTimeStep, timeStepIncr) (WallFieldSaverPrmsPy.cpp:287-330):
    fieldName "Force" or "Position" (CWall::getVectorFieldFunction, Wall.cpp:72-87)
    of one wall or a list of walls; fileFormat "RAW_SERIES" appends one line per
    save, each wall's vector in scientific notation with 10 digits followed by two
    spaces (VectorWallFieldMaster::writeAsRAW_SERIES, WallFieldMaster.cpp:24-44)."""

    def __init__(self, wallName, fieldName, fileName, fileFormat, beginTimeStep, endTimeStep,
                 timeStepIncr):
        FieldSaverPrms.__init__(self, fieldName, fileName, fileFormat, beginTimeStep, endTimeStep,
                                timeStepIncr)
        self.wallNames = [wallName] if isinstance(wallName, str) else list(wallName)
        if fieldName not in ("Force", "Position"):
            raise ValueError("invalid name [ %s ] for wall vector field access function" % fieldName)
        if fileFormat != "RAW_SERIES":
            raise ValueError("wall fields are saved as RAW_SERIES only")

    def getWallNameList(self):
        return list(self.wallNames)


class InteractionScalarFieldSaverPrms(FieldSaverPrms):
    """InteractionScalarFieldSaverPrms(interactionName, fieldName, fileName,
    fileFormat, beginTimeStep, endTimeStep, timeStepIncr)
    (InteractionFieldSaverPrmsPy.cpp:32-52): fieldName "count" (bonds: every
    bond; RotFriction: touching pairs, RotFricInteraction.cpp:447-452;
    RotElastic: every pair) or "breaking_criterion" of a bonded group
    (CRotBondedInteraction::getCriterion, RotBondedInteraction.cpp:221-227);
    fileFormat "SUM" appends the sum over the group's interactions in list order,
    scientific notation with 10 digits (ScalarInteractionFieldMaster::writeAsSUM,
    InteractionFieldMaster.cpp:460-487)."""

    def __init__(self, interactionName, fieldName, fileName, fileFormat, beginTimeStep,
                 endTimeStep, timeStepIncr):
        FieldSaverPrms.__init__(self, fieldName, fileName, fileFormat, beginTimeStep, endTimeStep,
                                timeStepIncr)
        self.interactionName = interactionName
        if fieldName not in ("count", "breaking_criterion"):
            raise ValueError("interaction field %r is not supported" % fieldName)
        if fileFormat != "SUM":
            raise ValueError("interaction scalar fields are saved as SUM only")


def _sci10(x):
    return "%.10e" % x


class GravityPrms(InteractionPrms):
    """GravityPrms(name, acceleration) (InteractionParamsPy.cpp:917-925)."""

    def __init__(self, name, acceleration):
        self.name, self.acceleration = name, Vec3(acceleration)


class DampingPrms(InteractionPrms):
    """DampingPrms(type, name, viscosity, maxIterations=100) with type "Damping" or
    "RotDamping" (InteractionParamsPy.cpp:36-62, 977-997): velocity damping solved
    by fixed-point iteration (Damping.hpp:87-107, RotDamping.hpp:58-84)."""

    _KIND = {"Damping": 0, "RotDamping": 1}

    def __init__(self, type, name, viscosity, maxIterations=100):
        if type not in self._KIND:
            raise RuntimeError("Trying to setup Damping of unknown type: %s" % type)
        self.type, self.name = type, name
        self.viscosity, self.maxIterations = float(viscosity), int(maxIterations)


class LinDampingPrms(DampingPrms):
    """LinDampingPrms(name, viscosity, maxIterations=100) = DampingPrms("Damping", ...)."""

    def __init__(self, name, viscosity, maxIterations=100):
        DampingPrms.__init__(self, "Damping", name, viscosity, maxIterations)


class RotDampingPrms(DampingPrms):
    """RotDampingPrms(name, viscosity, maxIterations=100) = DampingPrms("RotDamping", ...)."""

    def __init__(self, name, viscosity, maxIterations=100):
        DampingPrms.__init__(self, "RotDamping", name, viscosity, maxIterations)


class LocalDampingPrms(InteractionPrms):
    """LocalDampingPrms(name, viscosity): each force component against the sign
    of the velocity component (LocalDamping.hpp:82-117)."""

    type = "LocalDamping"

    def __init__(self, name, viscosity):
        self.name, self.viscosity = name, float(viscosity)


class RotLocalDampingPrms(LocalDampingPrms):
    """RotLocalDampingPrms(name, viscosity): the same on moment and angular
    velocity (RotLocalDamping.hpp:82-117)."""

    type = "RotLocalDamping"


class NRotBondedWallPrms(InteractionPrms):
    """NRotBondedWallPrms(name, wallName, normalK, particleTag, tagMask=-1)
    (WallPrmsPy.cpp:40-75, 130-166): the wall's candidate particles with
    (tag & tagMask) == (particleTag & tagMask) are bonded to it, the others
    interact elastically (BWallInteractionGroup.hpp)."""

    def __init__(self, name, wallName, normalK, particleTag, tagMask=-1):
        self.name, self.wallName, self.normalK = name, wallName, float(normalK)
        self.particleTag, self.tagMask = int(particleTag), int(tagMask)


class NRotElasticWallPrms(InteractionPrms):
    """NRotElasticWallPrms(name, wallName, normalK) (WallPrmsPy.cpp)."""

    def __init__(self, name, wallName, normalK):
        self.name, self.wallName, self.normalK = name, wallName, float(normalK)


# --------------------------------------------------------------------------
# LsmMpi
# --------------------------------------------------------------------------
_PTYPES = {"NRotSphere": _native.ESP_PTYPE_NROT, "RotSphere": _native.ESP_PTYPE_ROT}


class LsmMpi:
    """Model driver with the reference's ``esys.lsm.LsmMpi`` interface
    (LsmMpiPy.cpp:196-1300) over one GPU-resident subdomain."""

    def __init__(self, numWorkerProcesses=1, mpiDimList=(1, 1, 1), device=0, stepsPerBatch=32):
        self._device = int(device)
        self._steps_per_batch = int(stepsPerBatch)
        self._ptype = None
        self._grid = None
        self._verlet = None
        self._dt = 0.0
        self._nts = 0
        self._do2d = False
        self._engine = None
        self._pre = []
        self._post = []
        self._checkpointer = None
        self._savers = []
        self._started_saving = False
        self._group_kinds = {}

    # the library this facade binds; the product path is the HIP engine
    def _open_native(self):
        return _native.load_hip_library()

    # ---- model setup ----
    def initNeighbourSearch(self, particleType, gridSpacing, verletDist):
        if particleType not in _PTYPES:
            raise ValueError("particle type %r not supported (NRotSphere, RotSphere)" % particleType)
        self._ptype = _PTYPES[particleType]
        self._grid = float(gridSpacing)
        self._verlet = float(verletDist)

    initNeighborSearch = initNeighbourSearch
    initVerletModel = initNeighbourSearch

    def getParticleType(self):
        return {v: k for k, v in _PTYPES.items()}[self._ptype]

    def force2dComputations(self, do2d=True):
        if self._engine is not None:
            raise RuntimeError("force2dComputations must precede setSpatialDomain")
        self._do2d = bool(do2d)

    def setSpatialDomain(self, bbox, maxPt=None, circDimList=None):
        """setSpatialDomain(bbox) / setSpatialDomain(bbox, circDimList)
        (LsmMpiPy.cpp:1085-1100): circDimList = [x, y, z] circular flags;
        x only (the reference's ghost exchange is x-only, pp_array.hpp:433-536)."""
        if isinstance(maxPt, (list, tuple)) and circDimList is None:
            maxPt, circDimList = None, maxPt
        if maxPt is not None:
            bbox = BoundingBox(bbox, maxPt)
        if self._ptype is None:
            raise RuntimeError("initNeighbourSearch must be called before setSpatialDomain")
        if self._engine is not None:
            raise RuntimeError("spatial domain already set")
        circ = [bool(c) for c in (circDimList or (False, False, False))]
        if circ[1] or circ[2]:
            raise NotImplementedError("circular y or z boundaries are not supported "
                                      "(x only, as the reference's exchange)")
        self._circ = circ
        self._engine = _native.Engine(self._open_native(), self._ptype, self._grid, self._verlet,
                                      list(bbox.getMinPt()), list(bbox.getMaxPt()), self._dt,
                                      self._do2d, self._device, self._steps_per_batch,
                                      circular_x=circ[0])

    def _e(self):
        if self._engine is None:
            raise RuntimeError("setSpatialDomain has not been called")
        return self._engine

    @property
    def engine(self):
        return self._e()

    def setTimeStepSize(self, dt):
        self._dt = float(dt)
        if self._engine is not None:
            self._engine.set_time_step_size(dt)

    def getTimeStepSize(self):
        return self._dt

    def setNumTimeSteps(self, n):
        self._nts = int(n)

    def getNumTimeSteps(self):
        return self._nts

    def getTimeStep(self):
        return self._e().time_step()

    # ---- particles ----
    def createParticles(self, particles):
        ps = list(particles)
        if not ps:
            return
        ids = np.array([p.getId() for p in ps], np.int32)
        pos = np.array([list(p.getPosn()) for p in ps], np.float64)
        rad = np.array([p.getRadius() for p in ps], np.float64)
        tags = np.array([p.getTag() for p in ps], np.int32)
        masses = [p.getMass() for p in ps]
        if any(m is None for m in masses):
            mass = None if all(m is None for m in masses) else np.array(
                [(m if m is not None else (r ** 2 if self._do2d else r ** 3))
                 for m, r in zip(masses, rad)], np.float64)
        else:
            mass = np.array(masses, np.float64)
        self._e().add_particles(ids, pos, rad, mass, tags)

    def createParticle(self, particle):
        self.createParticles([particle])

    def readGeometry(self, fileName):
        """CLatticeMaster::readGeometry (LatticeMaster.hpp:29-78): the header's
        bounding box becomes the spatial domain unless one was set, then the
        particles (mass pi r^2, times 4/3 r in 3-D) and the connections of the
        .geo file are created (format: geometry.read_geo)."""
        from .geometry import read_geo
        g = read_geo(fileName)
        if self._engine is None:
            self.setSpatialDomain(BoundingBox(Vec3(g.info.bbox_min), Vec3(g.info.bbox_max)),
                                  circDimList=list(g.info.periodic))
        elif list(g.info.periodic) != list(getattr(self, "_circ", [False] * 3)):
            raise ValueError("the .geo file's periodic dimensions differ from the domain's")
        if len(g.ids):
            self.createParticleArrays(g.ids, g.pos, g.rad, mass=g.mass, tag=g.tags)
        if len(g.conn):
            self.createConnectionArrays(g.conn[:, 0], g.conn[:, 1], g.conn[:, 2])

    def createParticleArrays(self, ids, posn, radius, mass=None, tag=None, vel=None):
        """Bulk particle creation from arrays (same semantics as createParticles)."""
        self._e().add_particles(ids, posn, radius, mass, tag, vel)

    def getNumParticles(self):
        return self._e().num_particles()

    def setParticleDensity(self, tag, mask, Density):
        self._e().set_particle_density(tag, mask, Density)

    def setParticleNonDynamic(self, tag):
        self._e().set_particle_non_dynamic(tag)

    def setParticleNonRotational(self, tag):
        self._e().set_particle_non_rotational(tag)

    def setParticleVelocity(self, id, Velocity):
        self._e().set_particle_vel(id, list(Vec3(Velocity)))

    def setParticleAngularVelocity(self, id, angularVelocity):
        self._e().set_particle_ang_vel(id, list(Vec3(angularVelocity)))

    # ---- connections / groups ----
    def createConnections(self, connections):
        cs = list(connections)
        if not cs:
            return
        self._e().add_connections([c.getP1Id() for c in cs], [c.getP2Id() for c in cs],
                                  [c.getTag() for c in cs])

    def createConnectionArrays(self, id1, id2, tag=None):
        """Bulk connection creation from arrays (same semantics as createConnections)."""
        self._e().add_connections(id1, id2, tag)

    def createInteractionGroup(self, prms):
        e = self._e()
        if isinstance(prms, RotBondPrms):
            e.add_rot_bond_group(prms.name, prms._c(), prms.perBondStrengths)
        elif isinstance(prms, BrittleBeamSCPrms):
            e.add_brittle_beam_sc_group(prms.name, prms._c())
        elif isinstance(prms, NRotElasticPrms):
            e.add_nrot_elastic_group(prms.name, prms.normalK, prms.scaling)
            self._group_kinds[prms.name] = "nrot_elastic"
        elif isinstance(prms, RotElasticPrms):
            e.add_rot_elastic_group(prms.name, prms.normalK, prms.scaling)
            self._group_kinds[prms.name] = "rot_elastic"
        elif isinstance(prms, RotFrictionPrms):
            self._group_kinds[prms.name] = "friction"
            e.add_rot_friction_group(prms.name, {
                "k_n": prms.normalK, "mu_d": prms.dynamicMu, "mu_s": prms.staticMu,
                "k_s": prms.shearK, "scaling": int(prms.scaling), "rigid": int(prms.rigid),
                "meanR_scaling": int(prms.meanR_scaling)})
        elif isinstance(prms, GravityPrms):
            e.add_gravity(prms.name, list(prms.acceleration))
        elif isinstance(prms, DampingPrms):
            e.add_damping_group(DampingPrms._KIND[prms.type], prms.name, prms.viscosity,
                                prms.maxIterations)
        elif isinstance(prms, LocalDampingPrms):
            e.add_damping_group(3 if prms.type == "RotLocalDamping" else 2, prms.name,
                                prms.viscosity, 0)
        elif isinstance(prms, NRotElasticWallPrms):
            e.add_elastic_wall_group(prms.name, prms.wallName, prms.normalK)
        elif isinstance(prms, NRotBondedWallPrms):
            e.add_bonded_wall_group(prms.name, prms.wallName, prms.normalK, prms.particleTag,
                                    prms.tagMask)
        else:
            raise TypeError("interaction parameters %r are not supported" % type(prms).__name__)

    def createInteractionGroupTagged(self, prms, *tags):
        """createInteractionGroupTagged (LsmMpiPy.cpp:750-835, 1203-1206):
        (prms, tag1, mask1, tag2, mask2) for a RotFriction, NRotElastic or
        RotElastic group that takes only pairs whose tags match
        (pi_storage_ed_t.hpp, pi_storage_ne_t.hpp); (NRotElasticWallPrms, tag,
        mask) for an elastic wall acting only on particles with that tag
        (TaggedEWallInteractionGroup.hpp)."""
        if isinstance(prms, NRotElasticWallPrms):
            tag, mask = tags
            self.createInteractionGroup(prms)
            self._e().set_wall_group_tags(prms.name, tag, mask)
            return
        if not isinstance(prms, (RotFrictionPrms, NRotElasticPrms, RotElasticPrms)):
            raise TypeError("tagged groups are RotFriction, NRotElastic, RotElastic or "
                            "NRotElasticWall groups")
        tag1, mask1, tag2, mask2 = tags
        self.createInteractionGroup(prms)
        self._e().set_interaction_group_tags(prms.name, tag1, mask1, tag2, mask2)

    def createGravity(self, prms):
        self.createInteractionGroup(prms)

    def createDamping(self, prms):
        self.createInteractionGroup(prms)

    def createLocalDamping(self, prms):
        self.createInteractionGroup(prms)

    def createRotLocalDamping(self, prms):
        self.createInteractionGroup(prms)

    def createExclusion(self, interactionName1, interactionName2):
        self._e().add_exclusion(interactionName1, interactionName2)

    # ---- walls ----
    def createWall(self, name, posn, normal):
        self._e().add_wall(name, list(Vec3(posn)), list(Vec3(normal)))

    def moveWallBy(self, name, move):
        self._e().move_wall_by(name, list(Vec3(move)))

    def applyForceToWall(self, interactionName, force):
        """applyForceToWall (LsmMpiPy.cpp:1208): move the wall of the wall
        interaction group along the force until the group's wall force
        balances it (the servo of ServoWallLoaderRunnable)."""
        self._e().apply_force_to_wall(interactionName, list(force))

    def setWallStepMotion(self, name, move):
        """Move wall ``name`` by ``move`` at the start of every time step, as a
        pre-time-step Runnable calling moveWallBy would, without a Python call per step."""
        self._e().set_wall_step_motion(name, list(Vec3(move)))

    def getWallPosition(self, name):
        return Vec3(self._e().wall_position(name))

    def getWallForce(self, name):
        return Vec3(self._e().wall_force(name))

    # ---- runnables / stepping ----
    def addPreTimeStepRunnable(self, runnable):
        self._pre.append(runnable)

    def addPostTimeStepRunnable(self, runnable):
        self._post.append(runnable)

    # ---- restart checkpoints (LsmMpiPy.cpp:846-883) ----
    def createRestartCheckPointer(self, prms):
        """Restartable checkpoints in the reference's text format; a later call
        replaces the earlier checkpointer, as the reference's single controller does."""
        self._checkpointer = prms

    def loadCheckPoint(self, fileName):
        """Load <prefix>_t=<step>_0.txt (and the subdomain files it names) into a
        model whose groups and walls the script has created, before the first step."""
        self._e().load_checkpoint(fileName)

    # ---- field savers ----
    def createFieldSaver(self, prms):
        """LsmMpiPy::createFieldSaver (LsmMpiPy.cpp): wall vector fields and
        interaction scalar fields, written after the steps the parameters name
        (CLatticeMaster::runOneStep, LatticeMaster.cpp:2611-2632, 2672-2693)."""
        e = self._e()
        if isinstance(prms, WallVectorFieldSaverPrms):
            for w in prms.wallNames:
                e.wall_position(w)  # raises for an unknown wall
        elif isinstance(prms, InteractionScalarFieldSaverPrms):
            kind = self._group_kind(prms.interactionName)
            if prms.fieldName == "breaking_criterion":
                if kind != "bond":
                    raise ValueError("breaking_criterion needs a bonded interaction group")
                e.track_bond_criteria()
            elif kind == "nrot_elastic":
                raise NotImplementedError("count of an NRotElastic group is not supported")
        else:
            raise TypeError("field saver parameters %r are not supported" % type(prms).__name__)
        self._savers.append(prms)

    def _group_kind(self, name):
        e = self._e()
        if e.interaction_count(name) < 0:
            raise ValueError("unknown interaction group %r" % name)
        return self._group_kinds.get(name, "bond")

    def _write_field(self, s, t):
        e = self._e()
        if isinstance(s, WallVectorFieldSaverPrms):
            vals = [e.wall_force(w) if s.fieldName == "Force" else e.wall_position(w)
                    for w in s.wallNames]
            line = "".join(" ".join(_sci10(c) for c in v) + "  " for v in vals) + "\n"
        else:
            kind = self._group_kind(s.interactionName)
            if s.fieldName == "breaking_criterion":
                total = 0.0
                for c in e.bond_criteria(s.interactionName)[1]:
                    total += float(c)
            elif kind == "friction":
                total = float(np.count_nonzero(e.friction_state(s.interactionName)[1] & 2))
            else:
                total = float(e.interaction_count(s.interactionName))
            line = _sci10(0.0 + total) + "\n"
        with open(s.fileName, "a") as f:
            f.write(line)

    def _save_fields(self, t):
        for s in self._savers:
            if s.needSave(t):
                self._write_field(s, t)

    def _next_event(self, t):
        steps = []
        if self._checkpointer is not None:
            steps.append(self._checkpointer.nextCheckPoint(t))
        steps += [s.nextSave(t) for s in self._savers]
        steps = [x for x in steps if x is not None]
        return min(steps) if steps else None

    def _save_if_due(self, t):
        self._save_fields(t)
        cp = self._checkpointer
        if cp is not None and cp.isCheckPoint(t):
            self._e().save_checkpoint(cp.fileNamePrefix, self._nts, cp.precision)

    def _first_step_checkpoint(self):
        # CLatticeMaster::runOneStep saves the fields of step 0 before the first
        # search and the step-0 checkpoint after it (LatticeMaster.cpp:2611-2640)
        if not self._started_saving and self._e().time_step() == 0:
            self._started_saving = True
            self._save_if_due(0)

    def runTimeStep(self):
        self._first_step_checkpoint()
        for r in self._pre:
            r.run()
        self._e().run(1)
        self._save_if_due(self.getTimeStep())
        for r in self._post:
            r.run()

    def runTimeSteps(self, n):
        """n x runTimeStep; queued on the device in one call when no Python
        runnables are registered (split at checkpoint steps)."""
        n = int(n)
        if self._pre or self._post:
            for _ in range(n):
                self.runTimeStep()
            return
        if self._checkpointer is None and not self._savers:
            self._e().run(n)
            return
        self._first_step_checkpoint()
        while n > 0:
            t = self.getTimeStep()
            nxt = self._next_event(t)
            k = n if nxt is None else min(n, nxt - t)
            self._e().run(k)
            n -= k
            self._save_if_due(t + k)

    def run(self):
        remaining = self._nts - self.getTimeStep()
        if remaining > 0:
            self.runTimeSteps(remaining)

    # ---- state access ----
    def getParticleIds(self):
        return self._e().ids()

    def getParticleData(self, field):
        """Per-particle field ('pos', 'vel', 'angvel', 'quat', 'force', 'moment',
        'rad', 'mass', 'oldpos') as an array in ascending id order."""
        return self._e().field(field)

    def getParticlePosn(self, id):
        ids = self._e().ids()
        k = np.searchsorted(ids, id)
        if k >= len(ids) or ids[k] != id:
            raise KeyError(id)
        return Vec3(self._e().field("pos")[k])

    def getInteractionPairs(self, name):
        return self._e().interactions(name)

    def getNumInteractions(self, name):
        return self._e().interaction_count(name)

    def getBondBreaks(self):
        return self._e().bond_breaks()

    def trackPairEvents(self, name):
        """Log the interactions the dynamic group `name` creates and removes,
        from the first time step on (call before it)."""
        self._e().track_pair_events(name)

    def getPairEvents(self, name):
        """(step, id1, id2, kind) arrays in (step, id1, id2) order: kind +1 for
        an interaction the group created in that step's update, -1 for one it
        removed (rebuilt list, erased particle)."""
        return self._e().pair_events(name)

    def getParticleStress(self):
        """Per-particle stress sums of the last step (CParticle::m_sigma,
        Particle.cpp:391-400), shape (n, 3, 3), ascending id order; needs a
        BrittleBeamSC group.  Divide by the particle volume for CParticle::sigma."""
        return self._e().particle_stress()

    def getStats(self):
        return self._e().stats()

    def exit(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None
