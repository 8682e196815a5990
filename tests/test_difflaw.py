"""The laws of the differentiable engine (esys-particle_amd/csrc/esp_difflaw.h),
compiled for the host: on doubles they equal the oracle's laws bit for bit
(so the closed forms that pin the oracle pin them too); on dual numbers their
tangents equal central finite differences of the double evaluation; the
hand-written reverse mode of the bond law equals the dual-number Jacobian's
transpose applied to random output adjoints, on every branch of the law."""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "esys-particle_amd", "csrc")
pd = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(pd)


@pytest.fixture(scope="module")
def dl(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    so = str(tmp_path_factory.mktemp("dl") / "libdl.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I", CSRC,
                    os.path.join(HERE, "difflaw_harness.cpp"), "-o", so], check=True)
    lib = ctypes.CDLL(so)
    lib.dl_bond.argtypes = [pd, pd, pd]
    lib.dl_bond_tangent.argtypes = [pd, pd, ctypes.c_int, pd]
    lib.dl_bond_vjp.argtypes = [pd, pd, pd, pd]
    lib.dl_fric_vjp.argtypes = [pd, pd, ctypes.c_int, pd, pd]
    lib.dl_fric_vjp.restype = ctypes.c_int
    lib.dl_fric.argtypes = [pd, pd, ctypes.c_int, pd]
    lib.dl_fric.restype = ctypes.c_int
    lib.dl_fric_tangent.argtypes = [pd, pd, ctypes.c_int, ctypes.c_int, pd]
    lib.dl_fric_tangent.restype = ctypes.c_int
    return lib


def quat(rng, amax):
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    t = rng.uniform(0, amax)
    return np.array([math.cos(t / 2), *(math.sin(t / 2) * a)])


def bond_case(rng):
    r1, r2 = rng.uniform(0.2, 1.0, 2)
    d0 = rng.normal(size=3)
    d0 *= (r1 + r2) / np.linalg.norm(d0)
    x1 = rng.normal(size=3)
    x2 = x1 + d0 * rng.uniform(0.98, 1.02) + 0.02 * rng.normal(size=3)
    q1, q2 = quat(rng, 0.3), quat(rng, 0.3)
    inp = np.array([*x1, *x2, *q1, *q2])
    prm = np.array([r1, r2, 1e5, 3e4, 2e4, 4e4, *d0])
    return inp, prm


def test_bond_double_equals_oracle(dl, olib):
    rng = np.random.default_rng(1)
    for _ in range(300):
        inp, prm = bond_case(rng)
        out = np.zeros(13)
        dl.dl_bond(P(inp), P(prm), P(out))
        o_in = np.array([*inp[0:3], *inp[6:10], prm[0], *inp[3:6], *inp[10:14], prm[1]])
        o_b = np.array([*prm[2:6], *prm[6:9]])
        o = np.zeros(16)
        olib.law_rot_bond(P(o_in), P(o_b), P(o))
        assert out[0:6].tobytes() == o[0:6].tobytes()          # F1, M1
        assert out[6:9].tobytes() == o[9:12].tobytes()         # M2
        assert out[9:13].tobytes() == o[12:16].tobytes()       # nForce .. bMoment


def fric_case(rng):
    r1, r2 = rng.uniform(0.2, 1.0, 2)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    x1 = rng.normal(size=3)
    x2 = x1 + n * (r1 + r2) * rng.uniform(0.97, 0.999)
    v1, v2, w1, w2 = (rng.normal(size=3) for _ in range(4))
    ff = rng.normal(size=3) * rng.uniform(0, 20)
    inp = np.array([*x1, *x2, *v1, *v2, *w1, *w2, *ff])
    prm = np.array([r1, r2, 1e5, 3e4, 0.6, 1e-4])
    return inp, prm


def test_friction_double_equals_oracle(dl, olib):
    rng = np.random.default_rng(2)
    for _ in range(300):
        inp, prm = fric_case(rng)
        slip = int(rng.integers(0, 2))
        out = np.zeros(12)
        fl = dl.dl_fric(P(inp), P(prm), slip, P(out))
        o_in = np.array([*inp[0:3], *inp[6:9], *inp[12:15], prm[0], *inp[3:6], *inp[9:12],
                         *inp[15:18], prm[1]])
        o_prm = np.array([prm[2], prm[3], prm[4], prm[4], prm[5]])
        st = np.array([*inp[18:21], float(slip)])
        o = np.zeros(13)
        olib.law_rot_friction(P(o_in), P(o_prm), P(st), P(o))
        assert fl & 1 and o[12] == 1.0
        assert out[0:6].tobytes() == o[0:6].tobytes()   # -Fn + Ffric, M1
        assert out[6:9].tobytes() == o[9:12].tobytes()  # M2
        assert out[9:12].tobytes() == st[0:3].tobytes()  # updated spring
        assert ((fl >> 1) & 1) == int(st[3])             # slip flag


def _fd_check(f_val, f_tan, inp, prm, nin, nout, scale_fn):
    """Each tangent against central differences at three step sizes (relative
    1e-5, 1e-6, 1e-7): near a point of high curvature the truncation error of
    the larger steps dominates, rounding that of the smaller ones, so the best
    of the three is compared."""
    for k in range(nin):
        t = np.zeros(nout)
        f_tan(inp, prm, k, t)
        base = inp[k] if k < len(inp) else prm[2 + k - len(inp)]

        def ev(delta):
            i2, p2 = inp.copy(), prm.copy()
            if k < len(inp):
                i2[k] += delta
            else:
                p2[2 + k - len(inp)] += delta
            o = np.zeros(nout)
            f_val(i2, p2, o)
            return o
        errs = []
        for rel in (1e-5, 1e-6, 1e-7):
            h = rel * max(abs(base), 1e-3)
            fd = (ev(h) - ev(-h)) / (2 * h)
            errs.append(np.max(np.abs(t - fd)) / scale_fn(t, fd))
        assert min(errs) <= 1e-5, (k, errs)


def test_bond_tangents_equal_finite_differences(dl):
    rng = np.random.default_rng(3)
    for _ in range(20):
        inp, prm = bond_case(rng)
        _fd_check(lambda i, p, o: dl.dl_bond(P(i), P(p), P(o)),
                  lambda i, p, k, o: dl.dl_bond_tangent(P(i), P(p), k, P(o)),
                  inp, prm, 16, 13, lambda t, fd: max(np.abs(fd).max(), 1.0))


def test_friction_tangents_equal_finite_differences(dl):
    rng = np.random.default_rng(4)
    for _ in range(20):
        inp, prm = fric_case(rng)
        for slip in (0, 1):
            _fd_check(lambda i, p, o: dl.dl_fric(P(i), P(p), slip, P(o)),
                      lambda i, p, k, o: dl.dl_fric_tangent(P(i), P(p), slip, k, P(o)),
                      inp, prm, 24, 12, lambda t, fd: max(np.abs(fd).max(), 1.0))


def _vjp_cases(rng):
    """Random bonds, and the special configurations whose branches random ones
    miss: both particles unrotated (no relative rotation: no twist, no
    bending), a bond along z (the bond frame's xy branch), a bond parallel to
    its creation vector (no lateral shear), pure twist (no bending) and pure
    bending about the bond frame's axis."""
    for _ in range(200):
        yield bond_case(rng)
    for _ in range(20):
        inp, prm = bond_case(rng)
        inp[6:14] = [1, 0, 0, 0, 1, 0, 0, 0]
        yield inp, prm
    for _ in range(20):
        inp, prm = bond_case(rng)
        inp[3:6] = inp[0:3] + np.array([0.0, 0.0, prm[0] + prm[1]]) * rng.uniform(0.99, 1.01)
        prm[6:9] = [0.0, 0.0, prm[0] + prm[1]]
        yield inp, prm
    for _ in range(20):
        inp, prm = bond_case(rng)
        d = prm[6:9] / np.linalg.norm(prm[6:9])
        inp[3:6] = inp[0:3] + d * (prm[0] + prm[1]) * rng.uniform(0.99, 1.01)
        inp[6:10] = [1, 0, 0, 0]
        yield inp, prm
    for _ in range(20):
        inp, prm = bond_case(rng)
        inp[6:10] = [1, 0, 0, 0]
        d = (inp[3:6] - inp[0:3]) / np.linalg.norm(inp[3:6] - inp[0:3])
        t = rng.uniform(-0.3, 0.3)
        inp[10:14] = [math.cos(t / 2), *(math.sin(t / 2) * d)]
        yield inp, prm


def test_bond_reverse_mode_equals_dual_jacobian(dl):
    rng = np.random.default_rng(11)
    worst = 0.0
    for inp, prm in _vjp_cases(rng):
        lam = rng.normal(size=9)
        cols = np.zeros((16, 13))
        for k in range(16):
            dl.dl_bond_tangent(P(inp), P(prm), k, P(cols[k]))
        want = cols[:, :9] @ lam
        got = np.zeros(16)
        dl.dl_bond_vjp(P(inp), P(prm), P(lam), P(got))
        # each input's error against the magnitude of the terms its sum adds
        scale = np.abs(cols[:, :9]) @ np.abs(lam) + 1e-300
        err = np.max(np.abs(got - want) / scale)
        worst = max(worst, err)
        assert err <= 1e-9, (err, inp, prm)
    print("worst relative difference %.3g" % worst)


def test_bond_reverse_mode_planted_mistake_fails(dl):
    """The comparison above catches a plausible reverse-mode slip: dropping
    the kr column's contribution (g_kr) or negating the x1 column."""
    rng = np.random.default_rng(12)
    inp, prm = bond_case(rng)
    lam = rng.normal(size=9)
    cols = np.zeros((16, 13))
    for k in range(16):
        dl.dl_bond_tangent(P(inp), P(prm), k, P(cols[k]))
    want = cols[:, :9] @ lam
    got = np.zeros(16)
    dl.dl_bond_vjp(P(inp), P(prm), P(lam), P(got))
    scale = np.abs(cols[:, :9]) @ np.abs(lam) + 1e-300
    assert np.max(np.abs(got - want) / scale) <= 1e-9
    for bad in (lambda g: np.r_[g[:14], 0.0, g[15]], lambda g: np.r_[-g[:3], g[3:]]):
        assert np.max(np.abs(bad(got) - want) / scale) > 1e-3


def test_friction_reverse_mode_equals_dual_jacobian(dl):
    """The reverse mode of the friction law against the dual-number Jacobian
    (outputs f1, m1, m2 and the history written), on both slip branches; the
    dual inputs' order is (x1, x2, v1, v2, w1, w2, history, k, ks, mu)."""
    rng = np.random.default_rng(13)
    done = {0: 0, 1: 0}
    for it in range(400):
        inp, prm = fric_case(rng)
        if it % 2:  # a shear spring beyond the Coulomb limit: the slip branch
            inp[18:21] *= rng.uniform(1e2, 1e4)
        for slip in (0, 1):
            cols = np.zeros((24, 12))
            for k in range(24):
                f = dl.dl_fric_tangent(P(inp), P(prm), slip, k, P(cols[k]))
            if not f & 1:
                continue
            lam = rng.normal(size=12)
            want = cols @ lam
            got = np.zeros(24)
            assert dl.dl_fric_vjp(P(inp), P(prm), slip, P(lam), P(got)) == 1
            scale = np.abs(cols) @ np.abs(lam) + 1e-300
            err = np.max(np.abs(got - want) / scale)
            assert err <= 1e-9, (err, slip, inp, prm)
            done[(f >> 1) & 1] += 1
    assert done[0] > 20 and done[1] > 20  # both slip outcomes covered
