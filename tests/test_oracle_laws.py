"""Pins the oracle's interaction laws (oracle/esp_oracle.c) with what the
mathematics fixes: special configurations with hand-derived closed forms,
rigid-rotation covariance, angular-momentum balance, Newton's third law, and
beam theory for the scaled bond constants."""
import ctypes
import math

import numpy as np

pd = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(pd)


def quat_from_axis(axis, ang):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    return np.array([math.cos(ang / 2), *(math.sin(ang / 2) * axis)])


def qmul(a, b):
    s = a[0] * b[0] - a[1:] @ b[1:]
    v = a[0] * b[1:] + b[0] * a[1:] + np.cross(a[1:], b[1:])
    return np.array([s, *v])


def rotmat(q):
    s, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - s * z), 2 * (x * z + s * y)],
                     [2 * (x * y + s * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s * x)],
                     [2 * (x * z - s * y), 2 * (y * z + s * x), 1 - 2 * (x * x + y * y)]])


def bond(olib, p1, q1, r1, p2, q2, r2, kr, ks, kt, kb, D0):
    inp = np.array([*p1, *q1, r1, *p2, *q2, r2], float)
    b = np.array([kr, ks, kt, kb, *D0], float)
    out = np.zeros(16)
    olib.law_rot_bond(P(inp), P(b), P(out))
    return {"F1": out[0:3], "M1": out[3:6], "F2": out[6:9], "M2": out[9:12],
            "nForce": out[12], "shForce": out[13], "tMoment": out[14], "bMoment": out[15]}


I4 = np.array([1.0, 0, 0, 0])


def test_bond_zero_deformation(olib):
    o = bond(olib, [0, 0, 0], I4, 0.5, [1.0, 0, 0], I4, 0.5, 1e5, 3e4, 3e4, 3e4, [1.0, 0, 0])
    for k in ("F1", "M1", "F2", "M2"):
        np.testing.assert_allclose(o[k], 0, atol=1e-10)


def test_bond_pure_stretch(olib):
    """Axial stretch by 0.1 along the initial axis: Fr = kr (|d| - (r1+r2)) d/|d|
    acts on particle 1 towards particle 2; no shear, no moments."""
    kr = 2.5e4
    o = bond(olib, [0, 0, 0], I4, 0.5, [1.1, 0, 0], I4, 0.5, kr, 3e4, 3e4, 3e4, [1.0, 0, 0])
    np.testing.assert_allclose(o["F1"], [kr * 0.1, 0, 0], rtol=1e-12)
    np.testing.assert_array_equal(o["F2"], -o["F1"])
    np.testing.assert_allclose(o["M1"], 0, atol=1e-9)
    np.testing.assert_allclose(o["M2"], 0, atol=1e-9)
    assert o["nForce"] > 0 and abs(o["nForce"] - kr * 0.1) < 1e-8
    assert o["shForce"] < 1e-9


def test_bond_pure_compression_sign(olib):
    o = bond(olib, [0, 0, 0], I4, 0.5, [0.9, 0, 0], I4, 0.5, 1e4, 0, 0, 0, [1.0, 0, 0])
    np.testing.assert_allclose(o["F1"], [-1e4 * 0.1, 0, 0], rtol=1e-12)
    assert o["nForce"] < 0  # negative in compression (ARotBondedInteraction.cpp:436)


def test_bond_pure_twist(olib):
    """Bond along z, particle 2 twisted by psi about the bond axis: only the
    torsional moment kt*psi, +z on particle 1, -z on particle 2."""
    kt, psi = 7.0, 0.01
    o = bond(olib, [0, 0, 0], I4, 0.5, [0, 0, 1.0], quat_from_axis([0, 0, 1], psi), 0.5,
             1e4, 1e3, kt, 1e3, [0, 0, 1.0])
    np.testing.assert_allclose(o["M1"], [0, 0, kt * psi], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(o["M2"], [0, 0, -kt * psi], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(o["F1"], 0, atol=1e-9)
    assert abs(o["tMoment"] - kt * psi) < 1e-12


def test_bond_pure_bending(olib):
    """Bond along z, particle 2 rotated by theta about x (Wang et al. 2006 as
    coded: bending moment 0.5 kb theta, shear 0.5 ks theta r0 with its couple
    shared by eq_rad1/eq_rad2)."""
    kr, ks, kt, kb, th = 1e4, 3e3, 2e3, 5e3, 0.02
    r1, r2 = 0.4, 0.6
    o = bond(olib, [0, 0, 0], I4, r1, [0, 0, 1.0], quat_from_axis([1, 0, 0], th), r2,
             kr, ks, kt, kb, [0, 0, 1.0])
    r0 = r1 + r2
    eq1, eq2 = r1 / r0, r2 / r0
    np.testing.assert_allclose(o["F1"], [0, 0.5 * ks * th * r0, 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(o["M1"], [0.5 * kb * th - 0.5 * ks * th * r0 * eq1, 0, 0],
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(o["M2"], [-0.5 * ks * th * r0 * eq2 - 0.5 * kb * th, 0, 0],
                               rtol=1e-9, atol=1e-12)
    assert abs(o["bMoment"] - 0.5 * kb * th) < 1e-9


def _random_bond_config(rng):
    r1, r2 = rng.uniform(0.2, 1.0, 2)
    d0 = rng.normal(size=3)
    d0 *= (r1 + r2) * rng.uniform(0.995, 1.01) / np.linalg.norm(d0)
    p1 = rng.normal(size=3)
    q1 = quat_from_axis(rng.normal(size=3), rng.uniform(0, 0.05))
    q2 = quat_from_axis(rng.normal(size=3), rng.uniform(0, 0.05))
    # current relative vector: small perturbation of the body-frame D0 rotated by q1
    rel = rotmat(q1) @ (d0 * rng.uniform(0.99, 1.01) + rng.normal(size=3) * 0.01)
    return p1, q1, r1, p1 + rel, q2, r2, d0


def test_bond_rigid_rotation_covariance(olib):
    """Rotating both particles rigidly (positions and orientations) rotates the
    bond's forces and moments with them: F(R x, qR q) = R F(x, q)."""
    rng = np.random.default_rng(21)
    for _ in range(200):
        p1, q1, r1, p2, q2, r2, d0 = _random_bond_config(rng)
        k = (1e5, 3e4, 3e4, 3e4)
        o = bond(olib, p1, q1, r1, p2, q2, r2, *k, d0)
        qR = quat_from_axis(rng.normal(size=3), rng.uniform(0, 2 * math.pi))
        R = rotmat(qR)
        o2 = bond(olib, R @ p1, qmul(qR, q1), r1, R @ p2, qmul(qR, q2), r2, *k, d0)
        scale = max(np.abs(o["F1"]).max(), np.abs(o["M1"]).max(), 1e-3)
        for key in ("F1", "M1", "M2"):
            np.testing.assert_allclose(o2[key], R @ o[key], atol=1e-9 * scale)


def test_bond_angular_momentum_balance(olib):
    """An internal interaction exerts no net torque: M1 + M2 - (x2 - x1) x F1 = 0."""
    rng = np.random.default_rng(22)
    for _ in range(200):
        p1, q1, r1, p2, q2, r2, d0 = _random_bond_config(rng)
        o = bond(olib, p1, q1, r1, p2, q2, r2, 1e5, 3e4, 3e4, 3e4, d0)
        net = o["M1"] + o["M2"] - np.cross(p2 - p1, o["F1"])
        scale = np.abs(o["M1"]).max() + np.abs(o["M2"]).max() + np.abs(o["F1"]).max()
        np.testing.assert_allclose(net, 0, atol=1e-10 * scale + 1e-12)
        np.testing.assert_array_equal(o["F2"], -o["F1"])


def test_bond_constants_beam_theory(olib):
    """BrittleBeamPrms -> scaled bond constants equal the textbook beam values for
    a cylinder of radius R = d/2 and length L = d: EA/L, GJ/L, EI/L with A = pi R^2,
    J = pi R^4 / 2, I = pi R^4 / 4; strengths sigma A, tau A, tau J/R, sigma I/R."""
    E, nu, c, tan_phi = 2.0e5, 0.25, 50.0, 1.0
    G = E / (2 * (1 + nu))
    kr, ks, kt, kb = math.pi * E, math.pi * G, math.pi * G, math.pi * E
    mx = [math.pi * c / tan_phi, math.pi * c, math.pi * c, math.pi * c / tan_phi]
    p1, p2 = np.array([0.0, 0, 0]), np.array([0.3, 0.4, 0.0])  # d = 0.5
    inp = np.array([*p1, 0.3, *p2, 0.2])
    prm = np.array([kr, ks, kt, kb, *mx, 1, 1, 0], float)
    out = np.zeros(11)
    olib.bond_constants(P(inp), P(prm), P(out))
    d = 0.5
    R, L = d / 2, d
    A, J, I = math.pi * R ** 2, math.pi * R ** 4 / 2, math.pi * R ** 4 / 4
    sigma, tau = c / tan_phi, c
    np.testing.assert_allclose(out[:8], [E * A / L, G * A / L, G * J / L, E * I / L,
                                         sigma * A, tau * A, tau * J / R, sigma * I / R], rtol=1e-14)
    np.testing.assert_allclose(out[8:], p2 - p1, rtol=0)


def fric(olib, p1, v1, w1, r1, p2, v2, w2, r2, k, ks, mud, mus, dt, state):
    inp = np.array([*p1, *v1, *w1, r1, *p2, *v2, *w2, r2], float)
    prm = np.array([k, ks, mud, mus, dt], float)
    out = np.zeros(13)
    olib.law_rot_friction(P(inp), P(prm), P(state), P(out))
    return out


def test_friction_stick_then_slip(olib):
    """Constant tangential sliding velocity v: the shear spring grows by ks v dt
    per step while |Ffric| <= mu_s |Fn|, then saturates at mu_d |Fn|."""
    k, ks, mud, mus, dt, v = 100.0, 50.0, 0.4, 0.6, 1e-3, 2.0
    p1, p2 = [0, 0, 0], [0.95, 0, 0]  # overlap 0.05
    Fn = k * 0.05
    st = np.zeros(4)
    n_stick = 0
    for n in range(1, 200):
        out = fric(olib, p1, [0, 0, 0], [0, 0, 0], 0.5, p2, [0, v, 0], [0, 0, 0], 0.5,
                   k, ks, mud, mus, dt, st)
        if st[3] == 0:
            n_stick = n
            assert abs(st[1] - n * ks * v * dt) < 1e-12 * n
        else:
            assert abs(np.linalg.norm(st[:3]) - mud * Fn) < 1e-12
            break
    assert n_stick == int(mus * Fn / (ks * v * dt))  # last sticking step
    # normal force on particle 1 points away from particle 2; friction drags it along +y
    assert out[0] < 0 and out[1] > 0


def test_friction_no_contact_clears_history(olib):
    st = np.array([1.0, 2.0, 3.0, 1.0])
    out = fric(olib, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0.5, [1.2, 0, 0], [0, 0, 0], [0, 0, 0], 0.5,
               100, 50, 0.4, 0.6, 1e-3, st)
    np.testing.assert_array_equal(st, 0)
    assert out[12] == 0


def test_friction_rolling_contact_point(olib):
    """A spinning particle 1 (w along z) slips at the contact point with
    velocity w x (c - x1): the spring grows along that direction."""
    st = np.zeros(4)
    w = 3.0
    fric(olib, [0, 0, 0], [0, 0, 0], [0, 0, w], 0.5, [0.98, 0, 0], [0, 0, 0], [0, 0, 0], 0.5,
         100, 50, 0.4, 0.6, 1e-3, st)
    # contact point c = x2 + (r2/(r1+r2)) (x1 - x2) = (0.49, 0, 0); v_c1 = (0, 0.49 w, 0)
    np.testing.assert_allclose(st[:3], [0, -50 * 0.49 * w * 1e-3, 0], rtol=1e-12)


def fric_rigid(olib, p1, v1, w1, r1, p2, v2, w2, r2, k, ks, mud, mus, dt, state):
    """The rigid=True law; state = (Ffric, slipping, touching)."""
    inp = np.array([*p1, *v1, *w1, r1, *p2, *v2, *w2, r2], float)
    prm = np.array([k, ks, mud, mus, dt], float)
    out = np.zeros(13)
    olib.law_rot_friction_rigid(P(inp), P(prm), P(state), P(out))
    return out


def _tangent(D, rng):
    t = np.cross(D, rng.normal(size=3))
    return t / np.linalg.norm(t)


def test_rigid_friction_turns_with_the_pair(olib):
    """A touching, sticking pair moving as a rigid body (v_i = W x (x_i - c),
    w_i = W) has no slip at the contact point, and the rigid form turns the
    stored shear force with the pair: Ffric + dt W x Ffric
    (RotFricInteraction.cpp:283-297: spin plus axis rotation add up to W).
    The simple form leaves the force where it was."""
    rng = np.random.default_rng(4)
    k, ks, mud, mus, dt = 100.0, 50.0, 0.4, 0.6, 1e-3
    for W in ([0, 0, 2.0], [1.5, 0, 0], [0.3, -0.7, 1.1]):
        W = np.array(W)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        p1 = rng.normal(size=3)
        p2 = p1 + 0.95 * n  # overlap 0.05: |Fn| = 5, mu_s |Fn| = 3
        c = 0.5 * (p1 + p2)
        v1, v2 = np.cross(W, p1 - c), np.cross(W, p2 - c)
        F0 = 0.8 * _tangent(p1 - p2, rng)
        st = np.array([*F0, 0.0, 1.0])
        fric_rigid(olib, p1, v1, W, 0.5, p2, v2, W, 0.5, k, ks, mud, mus, dt, st)
        assert st[3] == 0 and st[4] == 1
        np.testing.assert_allclose(st[:3], F0 + dt * np.cross(W, F0), rtol=0, atol=1e-13)
        assert np.linalg.norm(dt * np.cross(W, F0)) > 1e-4  # the turn is resolved
        ss = np.array([*F0, 0.0])
        fric(olib, p1, v1, W, 0.5, p2, v2, W, 0.5, k, ks, mud, mus, dt, ss)
        np.testing.assert_allclose(ss[:3], F0, rtol=0, atol=1e-13)


def test_rigid_friction_without_rotation_matches_simple_law(olib):
    """Without particle rotation and from a first touch (any stored force is
    dropped, RotFricInteraction.cpp:277-281) both forms take the tangential
    relative displacement: the same spring, the same slip force."""
    rng = np.random.default_rng(5)
    k, ks, mud, mus, dt = 100.0, 50.0, 0.4, 0.6, 1e-3
    for trial in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        p1 = rng.normal(size=3)
        p2 = p1 + 0.95 * n
        v1, v2 = rng.normal(size=3), rng.normal(size=3) * (1 + 200 * (trial % 2))
        st = np.array([*rng.normal(size=3), 0.0, 0.0])
        o_r = fric_rigid(olib, p1, v1, [0, 0, 0], 0.5, p2, v2, [0, 0, 0], 0.5, k, ks, mud, mus,
                         dt, st)
        ss = np.zeros(4)
        o_s = fric(olib, p1, v1, [0, 0, 0], 0.5, p2, v2, [0, 0, 0], 0.5, k, ks, mud, mus, dt, ss)
        assert st[3] == ss[3] and st[4] == 1
        np.testing.assert_allclose(st[:3], ss[:3], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(o_r, o_s, rtol=1e-12, atol=1e-14)


def test_rigid_friction_slipping_branches(olib):
    """A slipping pair: with a slip increment above 1e-8 the force is
    mu_d |Fn| along it; without one the pair sticks and keeps its force
    (RotFricInteraction.cpp:310-321)."""
    k, ks, mud, mus, dt = 100.0, 50.0, 0.4, 0.6, 1e-3
    p1, p2 = [0, 0, 0], [0.95, 0, 0]
    st = np.array([0.0, 1.0, 0.0, 1.0, 1.0])
    fric_rigid(olib, p1, [0, 0, 0], [0, 0, 0], 0.5, p2, [0, 0, 3.0], [0, 0, 0], 0.5, k, ks, mud,
               mus, dt, st)
    assert st[3] == 1
    np.testing.assert_allclose(st[:3], [0, 0, mud * k * 0.05], rtol=1e-14)
    held = np.array([0.3, -0.2, 0.1, 1.0, 1.0])
    st = held.copy()
    fric_rigid(olib, p1, [0, 0, 0], [0, 0, 0], 0.5, p2, [0, 0, 0], [0, 0, 0], 0.5, k, ks, mud, mus,
               dt, st)
    assert st[3] == 0 and st[4] == 1
    np.testing.assert_array_equal(st[:3], held[:3])


def test_rigid_friction_no_contact_and_balance(olib):
    st = np.array([1.0, 2.0, 3.0, 1.0, 1.0])
    out = fric_rigid(olib, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0.5, [1.2, 0, 0], [0, 0, 0],
                     [0, 0, 0], 0.5, 100, 50, 0.4, 0.6, 1e-3, st)
    np.testing.assert_array_equal(st, 0)
    assert out[12] == 0
    # Newton's third law and the angular momentum balance about the origin
    rng = np.random.default_rng(6)
    for _ in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        p1 = rng.normal(size=3)
        p2 = p1 + 0.93 * n
        st = np.array([*(0.5 * _tangent(p1 - p2, rng)), float(rng.integers(2)), 1.0])
        out = fric_rigid(olib, p1, rng.normal(size=3), rng.normal(size=3), 0.5, p2,
                         rng.normal(size=3), rng.normal(size=3), 0.5, 100, 50, 0.4, 0.6, 1e-3, st)
        np.testing.assert_allclose(out[0:3], -out[6:9], rtol=0, atol=1e-14)
        L = out[3:6] + out[9:12] + np.cross(p1, out[0:3]) + np.cross(p2, out[6:9])
        assert np.linalg.norm(L) <= 1e-12 * (1 + np.linalg.norm(out[0:3]))


def test_nrot_elastic_hooke(olib):
    out = np.zeros(7)
    p = np.array([0, 0, 0, 0.5, 0.9, 0, 0, 0.5])
    olib.law_nrot_elastic(P(p), 2.0, P(out))
    np.testing.assert_allclose(out[:3], [-0.2, 0, 0], rtol=1e-14)
    np.testing.assert_array_equal(out[3:6], -out[:3])
