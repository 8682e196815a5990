/*
 * ref_math_harness.cpp - C entry points over the reference's own Foundation
 * math (Vec3, Matrix3, Quaternion), compiled from the reference sources by
 * oracle/Makefile into oracle/_ref/libref_math.so.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/test_oracle_ref_math.py to pin the
 * oracle's restated arithmetic against the reference's compiled operators
 * (Foundation/vec3.hpp, Foundation/Matrix3.hpp, Foundation/Quaternion.hpp).
 * Quaternions are passed as {scalar, x, y, z}; matrices row-major.
 */
#include "Foundation/vec3.h"
#include "Foundation/Matrix3.h"
#include "Foundation/Quaternion.h"

static Vec3 v_in(const double* a) { return Vec3(a[0], a[1], a[2]); }
static void v_out(const Vec3& v, double* o) { o[0] = v.X(); o[1] = v.Y(); o[2] = v.Z(); }
static Quaternion q_in(const double* a) { return Quaternion(a[0], Vec3(a[1], a[2], a[3])); }
static void q_out(const Quaternion& q, double* o)
{
  o[0] = q.return_sca();
  v_out(q.return_vec(), o + 1);
}
static Matrix3 m_in(const double* a)
{
  double m[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = a[3 * i + j];
  return Matrix3(m);
}
static void m_out(const Matrix3& m, double* o)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[3 * i + j] = m(i, j);
}

extern "C" {
void ref_quat_mul(const double* a, const double* b, double* o) { q_out(q_in(a) * q_in(b), o); }
void ref_quat_inverse(const double* a, double* o) { q_out(q_in(a).inverse(), o); }
void ref_quat_to_matrix(const double* a, double* o) { m_out(q_in(a).to_matrix(), o); }
void ref_mat_vec(const double* m, const double* v, double* o) { v_out(m_in(m) * v_in(v), o); }
void ref_mat_trans(const double* m, double* o) { m_out(m_in(m).trans(), o); }
void ref_vec_cross(const double* a, const double* b, double* o) { v_out(cross(v_in(a), v_in(b)), o); }
double ref_vec_dot(const double* a, const double* b) { return dot(v_in(a), v_in(b)); }
double ref_vec_norm(const double* a) { return v_in(a).norm(); }
double ref_vec_norm2(const double* a) { return v_in(a).norm2(); }
/* (dt/2.0)*(Quaternion(0, w)*q) added to q, then scaled by 1/|q|: the quaternion
 * update of a rotational particle step, written with the reference's operators. */
void ref_quat_step(const double* q, const double* w, double dt, double* o)
{
  Quaternion quat = q_in(q);
  quat += (dt / 2.0) * (Quaternion(0, v_in(w)) * quat);
  q_out(quat, o);
}
}
