// esp_difflaw.h - the interaction laws of the differentiable engine, written
// once over a scalar type T: T = double gives the forward values (bitwise the
// laws of esp_math.h: the same expression trees, compiled without contraction),
// T = esp_dual carries one directional derivative along, so a law evaluated on
// duals seeded with a unit tangent returns one column of its Jacobian.  The
// adjoint kernels build the vector-Jacobian products of the laws from such
// columns (local forward mode inside a global reverse sweep).
//
// Usable from host C++ (tests compile it with g++) and from HIP device code.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EDL_HD __host__ __device__ __forceinline__
#else
#define EDL_HD inline
#endif

// ---------------------------------------------------------------------------
// dual numbers: value and one tangent
// ---------------------------------------------------------------------------
struct esp_dual {
  double v, d;
  esp_dual() = default;
  EDL_HD constexpr esp_dual(double a, double b = 0.0) : v(a), d(b) {}
};
EDL_HD esp_dual edl_dual(double v, double d = 0.0) { return esp_dual(v, d); }
EDL_HD double edl_val(double a) { return a; }
EDL_HD double edl_val(esp_dual a) { return a.v; }
EDL_HD double edl_tan(double) { return 0.0; }
EDL_HD double edl_tan(esp_dual a) { return a.d; }

EDL_HD esp_dual operator+(esp_dual a, esp_dual b) { return {a.v + b.v, a.d + b.d}; }
EDL_HD esp_dual operator-(esp_dual a, esp_dual b) { return {a.v - b.v, a.d - b.d}; }
EDL_HD esp_dual operator-(esp_dual a) { return {-a.v, -a.d}; }
EDL_HD esp_dual operator*(esp_dual a, esp_dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
EDL_HD esp_dual operator/(esp_dual a, esp_dual b)
{
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
EDL_HD esp_dual operator+(esp_dual a, double b) { return {a.v + b, a.d}; }
EDL_HD esp_dual operator+(double a, esp_dual b) { return {a + b.v, b.d}; }
EDL_HD esp_dual operator-(esp_dual a, double b) { return {a.v - b, a.d}; }
EDL_HD esp_dual operator-(double a, esp_dual b) { return {a - b.v, -b.d}; }
EDL_HD esp_dual operator*(esp_dual a, double b) { return {a.v * b, a.d * b}; }
EDL_HD esp_dual operator*(double a, esp_dual b) { return {a * b.v, a * b.d}; }
EDL_HD esp_dual operator/(esp_dual a, double b) { return {a.v / b, a.d / b}; }
EDL_HD esp_dual operator/(double a, esp_dual b)
{
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
EDL_HD bool operator<(esp_dual a, double b) { return a.v < b; }
EDL_HD bool operator>(esp_dual a, double b) { return a.v > b; }
EDL_HD bool operator<=(esp_dual a, double b) { return a.v <= b; }
EDL_HD bool operator>=(esp_dual a, double b) { return a.v >= b; }
EDL_HD bool operator==(esp_dual a, double b) { return a.v == b; }
EDL_HD bool operator!=(esp_dual a, double b) { return a.v != b; }
EDL_HD bool operator<(esp_dual a, esp_dual b) { return a.v < b.v; }
EDL_HD bool operator>(esp_dual a, esp_dual b) { return a.v > b.v; }
EDL_HD esp_dual sqrt(esp_dual a)
{
  const double s = sqrt(a.v);
  return {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
EDL_HD esp_dual acos(esp_dual a)
{
  const double s = sqrt(1.0 - a.v * a.v);
  return {acos(a.v), s > 0.0 ? -a.d / s : 0.0};
}
EDL_HD esp_dual fmin(esp_dual a, esp_dual b) { return a.v <= b.v ? a : b; }
EDL_HD esp_dual fabs(esp_dual a) { return a.v < 0.0 ? -a : a; }

// ---------------------------------------------------------------------------
// vectors, quaternions, matrices over T (association order of Foundation/
// vec3.hpp, Matrix3.hpp, Quaternion.hpp, as esp_math.h)
// ---------------------------------------------------------------------------
template <class T>
struct tv3 {
  T x, y, z;
};
template <class T>
struct tquat {
  T s;
  tv3<T> v;
};
template <class T>
struct tmat3 {
  T m[3][3];
};

template <class T>
EDL_HD tv3<T> tmk(T x, T y, T z)
{
  return tv3<T>{x, y, z};
}
template <class T>
EDL_HD tv3<T> tadd(tv3<T> a, tv3<T> b)
{
  return tmk<T>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T>
EDL_HD tv3<T> tsub(tv3<T> a, tv3<T> b)
{
  return tmk<T>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class T>
EDL_HD tv3<T> tneg(tv3<T> a)
{
  return tmk<T>(-a.x, -a.y, -a.z);
}
template <class T, class S>
EDL_HD tv3<T> tmul(tv3<T> a, S s)
{
  return tmk<T>(a.x * s, a.y * s, a.z * s);
}
template <class T, class S>
EDL_HD tv3<T> tdiv(tv3<T> a, S s)
{
  return tmk<T>(a.x / s, a.y / s, a.z / s);
}
template <class T>
EDL_HD T tdot(tv3<T> a, tv3<T> b)
{
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
EDL_HD tv3<T> tcross(tv3<T> a, tv3<T> b)
{
  return tmk<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T>
EDL_HD T tnorm(tv3<T> a)
{
  return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
template <class T>
EDL_HD tv3<T> tmmul(const tmat3<T>& m, tv3<T> v)
{
  return tmk<T>(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
                m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
                m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}
template <class T>
EDL_HD tv3<T> tmtmul(const tmat3<T>& m, tv3<T> v)
{
  return tmk<T>(m.m[0][0] * v.x + m.m[1][0] * v.y + m.m[2][0] * v.z,
                m.m[0][1] * v.x + m.m[1][1] * v.y + m.m[2][1] * v.z,
                m.m[0][2] * v.x + m.m[1][2] * v.y + m.m[2][2] * v.z);
}
template <class T>
EDL_HD tquat<T> tqmul(tquat<T> a, tquat<T> b)
{
  tquat<T> r;
  r.s = a.s * b.s - tdot(a.v, b.v);
  r.v = tadd(tadd(tmul(b.v, a.s), tmul(a.v, b.s)), tcross(a.v, b.v));
  return r;
}
template <class T>
EDL_HD tquat<T> tqinv(tquat<T> a)
{
  return tquat<T>{a.s, tneg(a.v)};
}
template <class T>
EDL_HD tmat3<T> tqmat(tquat<T> q)
{
  tmat3<T> r;
  const T s = q.s, x = q.v.x, y = q.v.y, z = q.v.z;
  r.m[0][0] = s * s + x * x - y * y - z * z;
  r.m[0][1] = 2 * (x * y + s * z);
  r.m[0][2] = 2 * (x * z - s * y);
  r.m[1][0] = 2 * (x * y - s * z);
  r.m[1][1] = s * s - x * x + y * y - z * z;
  r.m[1][2] = 2 * (y * z + s * x);
  r.m[2][0] = 2 * (x * z + s * y);
  r.m[2][1] = 2 * (y * z - s * x);
  r.m[2][2] = s * s - x * x - y * y + z * z;
  return r;
}

// calc_angle (ARotBondedInteraction.cpp:22-51)
template <class T>
EDL_HD T tcalc_angle(T s_in, T c_os)
{
  if (s_in > 0.0) return (c_os >= 1.0 || c_os <= -1.0) ? T(0.0) : acos(c_os);
  if (s_in == 0.0) return T(0.0);
  return (c_os >= 1.0 || c_os <= -1.0) ? T(0.0) : -acos(c_os);
}

template <class T>
struct TBondResult {
  tv3<T> f1, m1, m2;
  T nForce, shForce, tMoment, bMoment;
};

// ARotBondedInteraction::calcForces (ARotBondedInteraction.cpp:307-448), the
// expression tree of esp_rot_bond_law
template <class T>
EDL_HD void tbond_law(tv3<T> pos1, tquat<T> q1, double r1, tv3<T> pos2, tquat<T> q2, double r2,
                      T kr, T ks, T kt, T kb, tv3<T> D0, TBondResult<T>& o)
{
  const double HALF_SQRT_2 = 0.5 * sqrt(2.0);
  T s_fai = T(0.0), c_fai = T(0.0);
  T sita = T(0.0), pasi = T(0.0);
  const tmat3<T> mat0 = tqmat(q1);
  const tv3<T> rbp = tsub(pos2, pos1);
  const T rbpNorm = tnorm(rbp);
  const tv3<T> rb = tmmul(mat0, rbp);
  const T rbNorm = tnorm(rb);
  const double r_0Norm = r1 + r2;
  const tv3<T> delta_r = tdiv(tmul(rb, rbNorm - r_0Norm), rbNorm);
  const tv3<T> Fr = tmul(delta_r, kr);

  const T gama_cos = tdot(rb, D0) / (rbNorm * tnorm(D0));
  T gama = T(0.0);
  if (gama_cos < 1.0 && gama_cos > -1.0) gama = acos(gama_cos);

  const tv3<T> rb_0 = tcross(rb, D0);
  const tv3<T> rb_b0 = tcross(rb, rb_0);
  const T rb_b0Norm = tnorm(rb_b0);
  tv3<T> Fst = tmk<T>(T(0.0), T(0.0), T(0.0));
  if (rb_b0Norm != 0.0) Fst = tdiv(tmul(rb_b0, ks * r_0Norm * gama), rb_b0Norm);

  tv3<T> Mst = tmk<T>(T(0.0), T(0.0), T(0.0));
  const T rb_0Norm = tnorm(rb_0);
  if (rb_0Norm != 0.0) Mst = tdiv(tmul(rb_0, -tnorm(Fst)), rb_0Norm);

  const T m0 = HALF_SQRT_2 * sqrt((rbNorm + rb.z) / rbNorm);
  T m1 = -HALF_SQRT_2 * sqrt((rbNorm - rb.z) / rbNorm);
  T m2 = -m1;
  const T m3 = T(0.0);
  const T rbX2PlusRbY2 = rb.x * rb.x + rb.y * rb.y;
  if (rbX2PlusRbY2 != 0.0) {
    const T denomSqrt = sqrt(rbX2PlusRbY2);
    m1 = m1 * rb.y / denomSqrt;
    m2 = m2 * rb.x / denomSqrt;
  }

  const tquat<T> qm = {m0, {m1, m2, m3}};
  const tmat3<T> mat = tqmat(qm);  // applied transposed below (matTrans)
  const tquat<T> rp = tqmul(tqinv(q1), q2);
  const tquat<T> r = tqmul(tqmul(tqinv(qm), rp), qm);
  const T temp0 = r.s * r.s + r.v.z * r.v.z;
  if (temp0 == 0.0) {
    pasi = T(0.0);
  } else {
    const T sqrtTemp0 = sqrt(temp0);
    const T c_pasi2 = r.s / sqrtTemp0;
    const T s_pasi2 = r.v.z / sqrtTemp0;
    pasi = 2.0 * tcalc_angle(s_pasi2, c_pasi2);
  }
  const tv3<T> Mtr = tmtmul(mat, tmk<T>(T(0.0), T(0.0), kt * pasi));
  const T temp2 = r.v.x * r.v.x + r.v.y * r.v.y;
  const T temp3 = r.v.x * r.v.z + r.s * r.v.y;
  const T temp4 = r.v.y * r.v.z - r.s * r.v.x;

  tv3<T> Mbr = tmk<T>(T(0.0), T(0.0), T(0.0)), Fsr = Mbr, Msr = Mbr;
  if (temp2 != 0.0) {
    const T temp1 = temp0 - temp2;
    if (temp1 >= 1.0 || temp1 <= -1.0) sita = T(0.0);
    else sita = acos(temp1);
    if (temp0 == 0.0) {
      const T sqrtDenom = sqrt(temp2);
      c_fai = r.v.y / sqrtDenom;
      s_fai = -r.v.x / sqrtDenom;
    } else {
      const T sqrtDenom = sqrt(temp0 * temp2);
      c_fai = temp3 / sqrtDenom;
      s_fai = temp4 / sqrtDenom;
    }
    Mbr = tmtmul(mat, tmk<T>(-0.5 * kb * sita * s_fai, 0.5 * kb * sita * c_fai, T(0.0)));
    Fsr = tmtmul(mat, tmk<T>(-0.5 * ks * c_fai * sita * r_0Norm, -0.5 * ks * s_fai * sita * r_0Norm,
                             T(0.0)));
    Msr = tmtmul(mat, tmk<T>(0.5 * ks * s_fai * sita * r_0Norm, -0.5 * ks * c_fai * sita * r_0Norm,
                             T(0.0)));
  }

  const double eq_rad1 = r1 / (r1 + r2);
  const double eq_rad2 = r2 / (r1 + r2);
  o.f1 = tmtmul(mat0, tadd(tadd(Fr, Fst), Fsr));
  o.m1 = tmtmul(mat0, tadd(tadd(tadd(Mbr, Mtr), tmul(tmul(Msr, eq_rad1), rbpNorm)),
                           tmul(tmul(Mst, eq_rad1), rbpNorm)));
  o.m2 = tmtmul(mat0, tsub(tsub(tadd(tmul(tmul(Msr, eq_rad2), rbpNorm),
                                     tmul(tmul(Mst, eq_rad2), rbpNorm)),
                                Mbr),
                           Mtr));
  o.shForce = tnorm(tadd(Fst, Fsr));
  o.tMoment = tnorm(Mtr);
  o.bMoment = tnorm(Mbr);
  const tv3<T> D = tsub(pos1, pos2);
  const T dist = sqrt(tdot(D, D));
  o.nForce = ((dist < r_0Norm) ? -1.0 : 1.0) * tnorm(Fr);
}

// Size scaling of a rotational bond's stiffnesses (RotBondedInteraction.cpp:
// 122-161, meanR or min radius, 3-D; 2-D unscaled), as esp_bond_constants
// computes them from the creation distance |D0|.
template <class T>
EDL_HD void tbond_k(T kr, T ks, double kt, double kb, double dist, double r1, double r2,
                    int scaling, int meanR, int do2d, T& kro, T& kso, double& kto, double& kbo)
{
  double effR = 1.0, effL = 1.0, effA = 1.0, momentJ = 1.0, momentI = 1.0;
  if (scaling && !do2d) {
    effR = meanR ? 0.5 * dist : fmin(r1, r2);
    effL = dist;
    effA = effR * effR;
    momentJ = 0.5 * effR * effR * effR * effR;
    momentI = 0.5 * momentJ;
  }
  kro = kr * effA / effL;
  kso = ks * effA / effL;
  kbo = kb * momentI / effL;
  kto = kt * momentJ / effL;
}

// CRotFrictionInteraction constructor scaling (RotFricInteraction.cpp:103-146)
template <class T>
EDL_HD void tfriction_k(T k, T ks, int scaling, int meanR, double r1, double r2, int do2d, T& ko,
                        T& kso)
{
  double effR = 1.0, effL = 1.0, effA = 1.0;
  const double r0 = r1 + r2;
  if (scaling && !do2d) {
    effR = meanR ? 0.5 * r0 : fmin(r1, r2);
    effL = r0;
    effA = effR * effR;
  }
  ko = k * effA / effL;
  kso = ks * effA / effL;
}

// CRotFrictionInteraction::calcSimpleForces (RotFricInteraction.cpp:171-231),
// the expression tree of esp_rot_friction_law.  slipping is the history's flag
// of the last step on entry and this step's on exit; f1 = -Fn + Ffric on the
// lower-id particle.  Returns touching.
template <class T>
EDL_HD bool tfriction_law(tv3<T> pos1, tv3<T> vel1, tv3<T> w1, double r1, tv3<T> pos2,
                          tv3<T> vel2, tv3<T> w2, double r2, T k, T ks, T mu_d, T mu_s, double dt,
                          tv3<T>& ffric, bool& slipping, tv3<T>& f1, tv3<T>& m1, tv3<T>& m2)
{
  const T mu_current = slipping ? mu_d : mu_s;
  const tv3<T> D = tsub(pos1, pos2);
  T dist = tdot(D, D);
  const double eq_dist = r1 + r2;
  if (dist < (eq_dist * eq_dist)) {
    dist = sqrt(dist);
    const tv3<T> force = tmul(D, k * (dist - eq_dist) / dist);
    const tv3<T> pos = tadd(pos2, tmul(D, r2 / eq_dist));
    const tv3<T> vp1 = tadd(vel1, tcross(w1, tsub(pos, pos1)));
    const tv3<T> vp2 = tadd(vel2, tcross(w2, tsub(pos, pos2)));
    tv3<T> ds = tmul(tsub(vp2, vp1), dt);
    ds = tsub(ds, tmul(D, tdot(ds, D) / tdot(D, D)));
    ffric = tadd(ffric, tmul(ds, ks));
    const T FfricNorm = tnorm(ffric);
    const T forceNorm = tnorm(force);
    if (FfricNorm > forceNorm * mu_current) {
      ffric = tmul(ffric, (mu_d * forceNorm) / FfricNorm);
      slipping = true;
    } else {
      slipping = false;
    }
    m1 = tcross(tsub(pos, pos1), ffric);
    m2 = tcross(tsub(pos, pos2), tneg(ffric));
    f1 = tadd(tmul(force, -1.0), ffric);
    return true;
  }
  ffric = tmk<T>(T(0.0), T(0.0), T(0.0));
  slipping = false;
  f1 = ffric;
  m1 = ffric;
  m2 = ffric;
  return false;
}
