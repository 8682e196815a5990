"""Pins the oracle's Vec3/Matrix3/Quaternion restatement against the reference's
own Foundation classes compiled from source (oracle/_ref/libref_math.so, built
from Foundation/vec3.cpp, Matrix3.cpp, Quaternion.cpp by oracle/Makefile).
Bit-exact: both sides evaluate the same expression trees without FMA."""
import ctypes

import numpy as np
import pytest

pd = ctypes.POINTER(ctypes.c_double)


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(pd)


@pytest.fixture(scope="module")
def ref(orc):
    lib = orc.load_ref_math()
    if lib is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    return lib


def rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def test_quaternion_ops_bitwise(olib, ref):
    rng = np.random.default_rng(11)
    for _ in range(500):
        a, b = rand_quat(rng) * rng.uniform(0.5, 2), rand_quat(rng) * rng.uniform(0.5, 2)
        o1, o2 = np.zeros(4), np.zeros(4)
        olib.math_quat_mul(P(a), P(b), P(o1))
        ref.ref_quat_mul(P(a), P(b), P(o2))
        assert o1.tobytes() == o2.tobytes()
        olib.math_quat_inverse(P(a), P(o1))
        ref.ref_quat_inverse(P(a), P(o2))
        assert o1.tobytes() == o2.tobytes()
        m1, m2 = np.zeros(9), np.zeros(9)
        olib.math_quat_to_matrix(P(a), P(m1))
        ref.ref_quat_to_matrix(P(a), P(m2))
        assert m1.tobytes() == m2.tobytes()


def test_matrix_vector_ops_bitwise(olib, ref):
    rng = np.random.default_rng(12)
    for _ in range(500):
        m = rng.normal(size=9)
        v, w = rng.normal(size=3), rng.normal(size=3)
        o1, o2 = np.zeros(3), np.zeros(3)
        olib.math_mat_vec(P(m), P(v), P(o1))
        ref.ref_mat_vec(P(m), P(v), P(o2))
        assert o1.tobytes() == o2.tobytes()
        t1, t2 = np.zeros(9), np.zeros(9)
        olib.math_mat_trans(P(m), P(t1))
        ref.ref_mat_trans(P(m), P(t2))
        assert t1.tobytes() == t2.tobytes()
        olib.math_vec_cross(P(v), P(w), P(o1))
        ref.ref_vec_cross(P(v), P(w), P(o2))
        assert o1.tobytes() == o2.tobytes()
        assert olib.math_vec_dot(P(v), P(w)) == ref.ref_vec_dot(P(v), P(w))
        assert olib.math_vec_norm(P(v)) == ref.ref_vec_norm(P(v))


def test_quaternion_integration_step_bitwise(olib, ref):
    """(dt/2)(Quaternion(0,w) q) added to q (RotParticle.cpp:289-290)."""
    rng = np.random.default_rng(13)
    for _ in range(500):
        q = rand_quat(rng)
        w = rng.normal(size=3) * 10
        dt = 10 ** rng.uniform(-6, -2)
        o1, o2 = np.zeros(4), np.zeros(4)
        olib.math_quat_step(P(q), P(w), dt, P(o1))
        ref.ref_quat_step(P(q), P(w), dt, P(o2))
        assert o1.tobytes() == o2.tobytes()


def test_to_matrix_is_inverse_rotation(olib):
    """Quaternion::to_matrix returns the transpose of the rotation matrix of q
    (the matrix that takes world vectors into the body frame); the bond law's
    frame algebra depends on it."""
    rng = np.random.default_rng(14)
    for _ in range(100):
        q = rand_quat(rng)
        s, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - s * z), 2 * (x * z + s * y)],
                      [2 * (x * y + s * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s * x)],
                      [2 * (x * z - s * y), 2 * (y * z + s * x), 1 - 2 * (x * x + y * y)]])
        m = np.zeros(9)
        olib.math_quat_to_matrix(P(q), P(m))
        np.testing.assert_allclose(m.reshape(3, 3), R.T, atol=1e-14)
