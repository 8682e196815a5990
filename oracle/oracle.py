"""Python driver of the CPU oracle (test infrastructure only).

Binds oracle/liboracle.so (the C restatement of the reference's hot path) and,
when built, oracle/_ref/libref_math.so (the reference's own Foundation math).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_MATH_PATH = os.path.join(HERE, "_ref", "libref_math.so")

_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

ORACLE_EXTRA = {
    "get_group_list_order": (ctypes.c_int, [_vp, ctypes.c_char_p, _pi, _pi]),
    "get_force_scale": (ctypes.c_int, [_vp, _pd]),
    "law_rot_bond": (None, [_pd, _pd, _pd]),
    "law_rot_friction": (None, [_pd, _pd, _pd, _pd]),
    "law_rot_friction_rigid": (None, [_pd, _pd, _pd, _pd]),
    "law_nrot_elastic": (None, [_pd, ctypes.c_double, _pd]),
    "bond_constants": (None, [_pd, _pd, _pd]),
    "math_quat_mul": (None, [_pd, _pd, _pd]),
    "math_quat_inverse": (None, [_pd, _pd]),
    "math_quat_to_matrix": (None, [_pd, _pd]),
    "math_mat_vec": (None, [_pd, _pd, _pd]),
    "math_mat_trans": (None, [_pd, _pd]),
    "math_vec_cross": (None, [_pd, _pd, _pd]),
    "math_vec_dot": (ctypes.c_double, [_pd, _pd]),
    "math_vec_norm": (ctypes.c_double, [_pd]),
    "math_quat_step": (None, [_pd, _pd, ctypes.c_double, _pd]),
    "math_rot_particle_step": (None, [_pd, ctypes.c_double]),
}


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_LIB = None


def load(pkg):
    """Return the oracle as a NativeLib of package ``pkg`` (eso_ prefix)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_PATH):
            build()
        _LIB = pkg._native.NativeLib(ORACLE_PATH, "eso_", ORACLE_EXTRA)
    return _LIB


def load_ref_math():
    """The reference's compiled Foundation math, or None when not built here."""
    if not os.path.exists(REF_MATH_PATH):
        return None
    lib = ctypes.CDLL(REF_MATH_PATH)
    for n, (res, args) in {
        "ref_quat_mul": (None, [_pd, _pd, _pd]),
        "ref_quat_inverse": (None, [_pd, _pd]),
        "ref_quat_to_matrix": (None, [_pd, _pd]),
        "ref_mat_vec": (None, [_pd, _pd, _pd]),
        "ref_mat_trans": (None, [_pd, _pd]),
        "ref_vec_cross": (None, [_pd, _pd, _pd]),
        "ref_vec_dot": (ctypes.c_double, [_pd, _pd]),
        "ref_vec_norm": (ctypes.c_double, [_pd]),
        "ref_vec_norm2": (ctypes.c_double, [_pd]),
        "ref_quat_step": (None, [_pd, _pd, ctypes.c_double, _pd]),
    }.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    return lib


def oracle_lsm_class(pkg):
    """LsmMpi facade subclass bound to the oracle instead of the HIP engine."""

    class OracleLsmMpi(pkg.LsmMpi):
        def _open_native(self):
            return load(pkg)

    return OracleLsmMpi


def oracle_engine(pkg, *args, **kw):
    """pkg.Engine over the oracle library."""
    return pkg.Engine(load(pkg), *args, **kw)
