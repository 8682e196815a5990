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
        _lib.esp_geom_pack_steps.restype = ctypes.c_double
        _lib.esp_geom_pack_steps.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int,
                                             ctypes.c_ulonglong, ctypes.c_int, pd, pd]
        _lib.esp_geom_pack_px.restype = ctypes.c_double
        _lib.esp_geom_pack_px.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int, ctypes.c_ulonglong,
                                          ctypes.c_int, pd]
        _lib.esp_geom_rsa.restype = ctypes.c_int
        _lib.esp_geom_rsa.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int, ctypes.c_ulonglong,
                                      ctypes.c_longlong, pd]
        _lib.esp_geom_snap_px.restype = ctypes.c_double
        _lib.esp_geom_snap_px.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), pd, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double]
        _lib.esp_geom_snap.restype = ctypes.c_double
        _lib.esp_geom_snap.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), pd, ctypes.c_int,
                                       ctypes.c_double]
    return _lib


def dense_packing(rad, box, seed, max_sweeps=6000, periodic_x=False, progress=None):
    """Positions of spheres with radii ``rad`` packed by grow-and-relax in
    [0, box] (x circular with ``periodic_x``); returns (pos, max relative
    overlap).  With ``progress`` the sweeps run in rounds of 100, each reported
    (the result is the same)."""
    lib = _load()
    rad = np.ascontiguousarray(rad, np.float64)
    box = np.ascontiguousarray(box, np.float64)
    pos = np.zeros((len(rad), 3))
    pd = ctypes.POINTER(ctypes.c_double)
    state = np.zeros(3)
    done, ov = 0, 0.0
    chunk = 100 if progress else int(max_sweeps)
    while done < max_sweeps:
        k = min(chunk, int(max_sweeps) - done)
        ov = lib.esp_geom_pack_steps(len(rad), rad.ctypes.data_as(pd), box.ctypes.data_as(pd),
                                     int(bool(periodic_x)), ctypes.c_ulonglong(seed), k,
                                     pos.ctypes.data_as(pd), state.ctypes.data_as(pd))
        done += k
        if progress:
            progress("packing: %d sweeps, scale %.3f, overlap %.3g" % (done, state[0], ov))
        if state[2]:
            break
    return pos, ov


def rsa_packing(rad, box, seed, periodic_x=False, max_tries=100000):
    """Random sequential insertion of spheres with radii ``rad`` (in that order)
    into [0, box] (x periodic when ``periodic_x``): positions, or RuntimeError
    when a sphere finds no free place in ``max_tries`` attempts."""
    lib = _load()
    rad = np.ascontiguousarray(rad, np.float64)
    box = np.ascontiguousarray(box, np.float64)
    pos = np.zeros((len(rad), 3))
    pd = ctypes.POINTER(ctypes.c_double)
    k = lib.esp_geom_rsa(len(rad), rad.ctypes.data_as(pd), box.ctypes.data_as(pd),
                         int(bool(periodic_x)), ctypes.c_ulonglong(seed), int(max_tries),
                         pos.ctypes.data_as(pd))
    if k != len(rad):
        raise RuntimeError("random packing too dense: placed %d of %d" % (k, len(rad)))
    return pos


def snap_bonds(pos, rad, bonds, box, max_sweeps=20000, tol=1e-10, periodic_x=False,
               progress=None):
    """Move centres so that every bonded pair is at exact contact and no other
    pair overlaps (x circular with ``periodic_x``); returns (pos, max relative
    residual).  With ``progress`` the sweeps run in rounds of 100, each
    reported (the result is the same)."""
    if progress:
        done, res = 0, 0.0
        while done < max_sweeps:
            k = min(100, int(max_sweeps) - done)
            pos, res = snap_bonds(pos, rad, bonds, box, k, tol, periodic_x)
            done += k
            progress("relaxation: %d sweeps, residual %.3g" % (done, res))
            if res < tol:
                break
        return pos, res
    lib = _load()
    pd = ctypes.POINTER(ctypes.c_double)
    pos = np.ascontiguousarray(pos, np.float64).copy()
    rad = np.ascontiguousarray(rad, np.float64)
    b = np.ascontiguousarray(bonds, np.int32)
    box = np.ascontiguousarray(box, np.float64)
    res = lib.esp_geom_snap_px(len(rad), pos.ctypes.data_as(pd), rad.ctypes.data_as(pd), len(b),
                               b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               box.ctypes.data_as(pd), int(bool(periodic_x)), int(max_sweeps),
                               float(tol))
    return pos, res


# ---------------------------------------------------------------------------
# .geo files and distance bonds
# ---------------------------------------------------------------------------
class GeometryInfo:
    """Header of a .geo file (GeometryInfo::Impl::read, GeometryInfo.cpp:157-233)."""

    def __init__(self, version, bbox_min, bbox_max, periodic, is2d):
        self.version = version
        self.bbox_min, self.bbox_max = tuple(bbox_min), tuple(bbox_max)
        self.periodic = tuple(periodic)
        self.is2d = is2d

    def hasAnyPeriodicDimensions(self):
        return any(self.periodic)


class GeoFile:
    """A parsed .geo file: header, particles and connections as arrays."""

    def __init__(self, info, ids, tags, pos, rad, mass, conn):
        self.info = info
        self.ids, self.tags, self.pos, self.rad, self.mass = ids, tags, pos, rad, mass
        self.conn = conn  # (n, 3) rows: id1, id2, tag


def _stream_tokens(text):
    for tok in text.split():
        yield tok


def read_geo(path):
    """Parse a .geo file the way CLatticeMaster::readGeometry reads it
    (LatticeMaster.hpp:29-78, GeometryReader.cpp:33-160):

    - header "LSMGeometry <1.1|1.2|1.3>" (or a token starting "ESyS", read as
      1.2), "BoundingBox x0 y0 z0 x1 y1 z1", "PeriodicBoundaries px py pz",
      from 1.2 on "Dimension 2D|3D" (1.1 files are 2-D), in 1.3
      "GengeoVersion <version> <rev>";
    - the first "BeginParticles" token, the particle type (only "Simple" is
      accepted), the count, then "x y z r id tag" per particle; the mass is
      pi r^2, times 4/3 r in 3-D (GeometryReader.h:52-62);
    - the first "BeginConnect" token after them, the count, then "id1 id2 tag"."""
    with open(path) as f:
        toks = f.read().split()
    k = 0

    def nxt():
        nonlocal k
        if k >= len(toks):
            raise RuntimeError("unexpected end of geometry file %s" % path)
        k += 1
        return toks[k - 1]

    ftype = nxt()
    if ftype != "LSMGeometry" and not ftype[:4] == "ESyS":
        raise RuntimeError("Unrecognised file type %s expected LSMGeometry or ESyS-Particle." % ftype)
    version = float(nxt())
    if ftype == "LSMGeometry" and version not in (1.1, 1.2, 1.3):
        raise RuntimeError("Can only read version 1.1, 1.2 or 1.3 geometry files, this is version %g"
                           % version)
    if ftype[:4] == "ESyS":
        version = 1.2
    if nxt() != "BoundingBox":
        raise RuntimeError("Expected BoundingBox")
    bb = [float(nxt()) for _ in range(6)]
    if nxt() != "PeriodicBoundaries":
        raise RuntimeError("Expected PeriodicBoundaries")
    periodic = [int(nxt()) != 0 for _ in range(3)]
    is2d = True
    if version >= 1.2:
        if nxt() != "Dimension":
            raise RuntimeError("Expected 'Dimension'")
        is2d = nxt() in ("2D", "2d")
    if version == 1.3:
        if nxt() != "GengeoVersion":
            raise RuntimeError("Expected 'GengeoVersion'")
        nxt()
        nxt()
    info = GeometryInfo(version, bb[:3], bb[3:], periodic, is2d)
    while k < len(toks) and toks[k] != "BeginParticles":
        k += 1
    if k >= len(toks):
        raise RuntimeError("Could not find 'BeginParticles' token in stream.")
    k += 1
    ptype = nxt()
    if ptype != "Simple":
        raise RuntimeError("Unknown particle type " + ptype)
    n = int(nxt())
    rec = np.array(toks[k:k + 6 * n], dtype=object).reshape(n, 6) if n else np.zeros((0, 6), object)
    k += 6 * n
    pos = rec[:, :3].astype(np.float64)
    rad = rec[:, 3].astype(np.float64)
    ids = rec[:, 4].astype(np.int64).astype(np.int32)
    tags = rec[:, 5].astype(np.int64).astype(np.int32)
    mass = np.pi * rad * rad
    if not is2d:
        mass = mass * ((4.0 / 3.0) * rad)
    conn = np.zeros((0, 3), np.int32)
    while k < len(toks) and toks[k] != "BeginConnect":
        k += 1
    if k < len(toks):
        k += 1
        nc = int(nxt())
        conn = np.array(toks[k:k + 3 * nc], np.int64).reshape(nc, 3).astype(np.int32)
    return GeoFile(info, ids, tags, pos, rad, mass, conn)


class DistConnections:
    """DistConnections(maxDist, bondTag, pList): bonds between particles whose
    surfaces are at most maxDist apart, |x1 - x2|^2 <= (maxDist + r1 + r2)^2
    (DistConnections.hpp:139-206, DistConnectionsPy.cpp:150-220).  Iterates
    (id1 < id2, tag) connections in ascending (id1, id2) order; the reference
    emits the same set in its neighbour table's order.  pList: particles with
    getId/getPosn/getRadius, or a (ids, pos, rad) tuple of arrays."""

    def __init__(self, maxDist, bondTag, pList):
        self.maxDist, self.bondTag = float(maxDist), int(bondTag)
        if isinstance(pList, tuple) and len(pList) == 3:
            ids, pos, rad = (np.asarray(pList[0], np.int64), np.asarray(pList[1], np.float64),
                             np.asarray(pList[2], np.float64))
        else:
            ps = list(pList)
            ids = np.array([p.getId() for p in ps], np.int64)
            pos = np.array([list(p.getPosn()) for p in ps], np.float64).reshape(-1, 3)
            rad = np.array([p.getRadius() for p in ps], np.float64)
        self._pairs = self._find(ids, pos, rad)

    def _find(self, ids, pos, rad):
        from .specimens import _cell_neighbours
        if len(ids) < 2:
            return np.zeros((0, 2), np.int64)
        cand = _cell_neighbours(pos, 2.0 * float(rad.max()) + self.maxDist)
        a, b = cand[:, 0], cand[:, 1]
        d = pos[a] - pos[b]
        rs = self.maxDist + rad[a] + rad[b]
        keep = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) <= rs * rs
        i1, i2 = ids[a[keep]], ids[b[keep]]
        lo, hi = np.minimum(i1, i2), np.maximum(i1, i2)
        order = np.lexsort((hi, lo))
        return np.stack([lo[order], hi[order]], 1)

    def getIdPairs(self):
        return self._pairs.copy()

    def getNumConnections(self):
        return len(self._pairs)

    def __len__(self):
        return len(self._pairs)

    def __iter__(self):
        from .lsm import TaggedIdConnection
        for a, b in self._pairs:
            yield TaggedIdConnection(int(a), int(b), self.bondTag)
