"""Dense seeded sphere packing (csrc/geometry.cpp, built into libesp_geometry.so)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEOM_LIB = os.path.join(HERE, "libesp_geometry.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(GEOM_LIB):
            raise RuntimeError("libesp_geometry.so is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(GEOM_LIB)
        pd = ctypes.POINTER(ctypes.c_double)
        _lib.esp_geom_pack.restype = ctypes.c_double
        _lib.esp_geom_pack.argtypes = [ctypes.c_int, pd, pd, ctypes.c_ulonglong, ctypes.c_int, pd]
        _lib.esp_geom_snap.restype = ctypes.c_double
        _lib.esp_geom_snap.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), pd, ctypes.c_int,
                                       ctypes.c_double]
    return _lib


def dense_packing(rad, box, seed, max_sweeps=6000):
    """Positions of spheres with radii ``rad`` packed by grow-and-relax in
    [0, box]; returns (pos, max relative overlap)."""
    lib = _load()
    rad = np.ascontiguousarray(rad, np.float64)
    box = np.ascontiguousarray(box, np.float64)
    pos = np.zeros((len(rad), 3))
    pd = ctypes.POINTER(ctypes.c_double)
    ov = lib.esp_geom_pack(len(rad), rad.ctypes.data_as(pd), box.ctypes.data_as(pd),
                           ctypes.c_ulonglong(seed), int(max_sweeps), pos.ctypes.data_as(pd))
    return pos, ov


def snap_bonds(pos, rad, bonds, box, max_sweeps=20000, tol=1e-10):
    """Move centres so that every bonded pair is at exact contact and no other
    pair overlaps; returns (pos, max relative residual)."""
    lib = _load()
    pd = ctypes.POINTER(ctypes.c_double)
    pos = np.ascontiguousarray(pos, np.float64).copy()
    rad = np.ascontiguousarray(rad, np.float64)
    b = np.ascontiguousarray(bonds, np.int32)
    box = np.ascontiguousarray(box, np.float64)
    res = lib.esp_geom_snap(len(rad), pos.ctypes.data_as(pd), rad.ctypes.data_as(pd), len(b),
                            b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            box.ctypes.data_as(pd), int(max_sweeps), float(tol))
    return pos, res
