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

// ---------------------------------------------------------------------------
// Reverse mode of tbond_law: given the adjoints lf1, lm1, lm2 of (f1, m1, m2),
// the adjoints of the inputs (x1, x2, q1, q2, kr, ks), on the branch the
// forward evaluation takes (the same branches as tbond_law<double>; at a
// square root or arc cosine of zero slope the derivative is taken as zero, as
// the dual numbers do).  One forward pass keeping its intermediates, then the
// chain rule backwards through it: a few forward passes' work instead of one
// dual pass per input.
// ---------------------------------------------------------------------------
namespace edl_rev {
typedef tv3<double> V;
typedef tquat<double> Q;
typedef tmat3<double> M;

EDL_HD V vz() { return V{0.0, 0.0, 0.0}; }
EDL_HD void vacc(V& a, V b) { a = tadd(a, b); }
// out = m^T v: v's adjoint m lo, m's adjoint v (x) lo
EDL_HD V mt_adj(const M& m, V v, V lo, M& lm)
{
  lm.m[0][0] += v.x * lo.x; lm.m[0][1] += v.x * lo.y; lm.m[0][2] += v.x * lo.z;
  lm.m[1][0] += v.y * lo.x; lm.m[1][1] += v.y * lo.y; lm.m[1][2] += v.y * lo.z;
  lm.m[2][0] += v.z * lo.x; lm.m[2][1] += v.z * lo.y; lm.m[2][2] += v.z * lo.z;
  return tmmul(m, lo);
}
// out = m v: v's adjoint m^T lo, m's adjoint lo (x) v
EDL_HD V m_adj(const M& m, V v, V lo, M& lm)
{
  lm.m[0][0] += lo.x * v.x; lm.m[0][1] += lo.x * v.y; lm.m[0][2] += lo.x * v.z;
  lm.m[1][0] += lo.y * v.x; lm.m[1][1] += lo.y * v.y; lm.m[1][2] += lo.y * v.z;
  lm.m[2][0] += lo.z * v.x; lm.m[2][1] += lo.z * v.y; lm.m[2][2] += lo.z * v.z;
  return tmtmul(m, lo);
}
// adjoint of tqmat(q) for the matrix adjoint lm
EDL_HD Q qmat_adj(Q q, const M& L)
{
  const double s = q.s, x = q.v.x, y = q.v.y, z = q.v.z;
  const double (*l)[3] = L.m;
  Q g;
  g.s = 2.0 * (s * l[0][0] + z * l[0][1] - y * l[0][2] - z * l[1][0] + s * l[1][1] + x * l[1][2] +
               y * l[2][0] - x * l[2][1] + s * l[2][2]);
  g.v.x = 2.0 * (x * l[0][0] + y * l[0][1] + z * l[0][2] + y * l[1][0] - x * l[1][1] + s * l[1][2] +
                 z * l[2][0] - s * l[2][1] - x * l[2][2]);
  g.v.y = 2.0 * (-y * l[0][0] + x * l[0][1] - s * l[0][2] + x * l[1][0] + y * l[1][1] + z * l[1][2] +
                 s * l[2][0] + z * l[2][1] - y * l[2][2]);
  g.v.z = 2.0 * (-z * l[0][0] + s * l[0][1] + x * l[0][2] - s * l[1][0] - z * l[1][1] + y * l[1][2] +
                 x * l[2][0] + y * l[2][1] + z * l[2][2]);
  return g;
}
// c = a b: adjoints of a and b for the adjoint lc of c
EDL_HD void qmul_adj(Q a, Q b, Q lc, Q& la, Q& lb)
{
  la.s += lc.s * b.s + tdot(lc.v, b.v);
  la.v = tadd(la.v, tadd(tsub(tmul(lc.v, b.s), tmul(b.v, lc.s)), tcross(b.v, lc.v)));
  lb.s += lc.s * a.s + tdot(lc.v, a.v);
  lb.v = tadd(lb.v, tadd(tsub(tmul(lc.v, a.s), tmul(a.v, lc.s)), tcross(lc.v, a.v)));
}
EDL_HD double dsqrt_half(double s) { return s > 0.0 ? 0.5 / s : 0.0; }   // d sqrt(u)/du
EDL_HD double dacos(double c)
{
  const double s = sqrt(1.0 - c * c);
  return s > 0.0 ? -1.0 / s : 0.0;
}
}  // namespace edl_rev

EDL_HD void tbond_law_vjp(tv3<double> pos1, tquat<double> q1, double r1, tv3<double> pos2,
                          tquat<double> q2, double r2, double kr, double ks, double kt, double kb,
                          tv3<double> D0, tv3<double> lf1, tv3<double> lm1, tv3<double> lm2,
                          tv3<double>& g_pos1, tv3<double>& g_pos2, tquat<double>& g_q1,
                          tquat<double>& g_q2, double& g_kr, double& g_ks)
{
  using namespace edl_rev;
  typedef double T;
  const double HALF_SQRT_2 = 0.5 * sqrt(2.0);
  // ---- forward (tbond_law's expressions), intermediates kept ----
  const M mat0 = tqmat(q1);
  const V rbp = tsub(pos2, pos1);
  const T rbpNorm = tnorm(rbp);
  const V rb = tmmul(mat0, rbp);
  const T rbNorm = tnorm(rb);
  const double r0 = r1 + r2;
  const T sdr = rbNorm - r0;
  const V delta_r = tdiv(tmul(rb, sdr), rbNorm);
  const V Fr = tmul(delta_r, kr);
  const T D0n = tnorm(D0);
  const T gden = rbNorm * D0n;
  const T gama_cos = tdot(rb, D0) / gden;
  const bool g_in = gama_cos < 1.0 && gama_cos > -1.0;
  const T gama = g_in ? acos(gama_cos) : 0.0;
  const V rb_0 = tcross(rb, D0);
  const V rb_b0 = tcross(rb, rb_0);
  const T B = tnorm(rb_b0);
  const T gmul = ks * r0 * gama;
  V Fst = vz();
  if (B != 0.0) Fst = tdiv(tmul(rb_b0, gmul), B);
  const T R0 = tnorm(rb_0);
  const T FstN = tnorm(Fst);
  V Mst = vz();
  if (R0 != 0.0) Mst = tdiv(tmul(rb_0, -FstN), R0);
  const T a0 = (rbNorm + rb.z) / rbNorm, a1 = (rbNorm - rb.z) / rbNorm;
  const T sa0 = sqrt(a0), sa1 = sqrt(a1);
  const T m0 = HALF_SQRT_2 * sa0;
  const T m1a = -HALF_SQRT_2 * sa1;
  T m1 = m1a, m2 = -m1a;
  const T xy2 = rb.x * rb.x + rb.y * rb.y;
  T dS = 0.0;
  if (xy2 != 0.0) {
    dS = sqrt(xy2);
    m1 = m1a * rb.y / dS;
    m2 = (-m1a) * rb.x / dS;
  }
  const Q qm = {m0, {m1, m2, 0.0}};
  const M mat = tqmat(qm);
  const Q iq1 = tqinv(q1), iqm = tqinv(qm);
  const Q rp = tqmul(iq1, q2);
  const Q u = tqmul(iqm, rp);
  const Q r = tqmul(u, qm);
  const T temp0 = r.s * r.s + r.v.z * r.v.z;
  T pasi = 0.0, sq0 = 0.0, c_p = 0.0, dang = 0.0;
  if (temp0 != 0.0) {
    sq0 = sqrt(temp0);
    c_p = r.s / sq0;
    const T s_p = r.v.z / sq0;
    const bool out = c_p >= 1.0 || c_p <= -1.0;
    if (s_p > 0.0) {
      pasi = 2.0 * (out ? 0.0 : acos(c_p));
      dang = out ? 0.0 : dacos(c_p);
    } else if (s_p < 0.0) {
      pasi = 2.0 * (out ? 0.0 : -acos(c_p));
      dang = out ? 0.0 : -dacos(c_p);
    }
  }
  const V mt = tmk<T>(0.0, 0.0, kt * pasi);
  const V Mtr = tmtmul(mat, mt);
  const T temp2 = r.v.x * r.v.x + r.v.y * r.v.y;
  const T temp3 = r.v.x * r.v.z + r.s * r.v.y;
  const T temp4 = r.v.y * r.v.z - r.s * r.v.x;
  V Mbr = vz(), Fsr = vz(), Msr = vz(), mb = vz(), fs = vz(), ms = vz();
  T sita = 0.0, c_fai = 0.0, s_fai = 0.0, sD = 0.0, temp1 = 0.0;
  const bool bend = temp2 != 0.0;
  if (bend) {
    temp1 = temp0 - temp2;
    if (!(temp1 >= 1.0 || temp1 <= -1.0)) sita = acos(temp1);
    if (temp0 == 0.0) {
      sD = sqrt(temp2);
      c_fai = r.v.y / sD;
      s_fai = -r.v.x / sD;
    } else {
      sD = sqrt(temp0 * temp2);
      c_fai = temp3 / sD;
      s_fai = temp4 / sD;
    }
    mb = tmk<T>(-0.5 * kb * sita * s_fai, 0.5 * kb * sita * c_fai, 0.0);
    fs = tmk<T>(-0.5 * ks * c_fai * sita * r0, -0.5 * ks * s_fai * sita * r0, 0.0);
    ms = tmk<T>(0.5 * ks * s_fai * sita * r0, -0.5 * ks * c_fai * sita * r0, 0.0);
    Mbr = tmtmul(mat, mb);
    Fsr = tmtmul(mat, fs);
    Msr = tmtmul(mat, ms);
  }
  const double eq1 = r1 / (r1 + r2), eq2 = r2 / (r1 + r2);
  const V A = tadd(tadd(Fr, Fst), Fsr);
  const V Bv = tadd(tadd(tadd(Mbr, Mtr), tmul(tmul(Msr, eq1), rbpNorm)), tmul(tmul(Mst, eq1), rbpNorm));
  const V C = tsub(tsub(tadd(tmul(tmul(Msr, eq2), rbpNorm), tmul(tmul(Mst, eq2), rbpNorm)), Mbr), Mtr);

  // ---- reverse ----
  M lmat0 = {}, lmat = {};
  const V lA = mt_adj(mat0, A, lf1, lmat0);
  const V lB = mt_adj(mat0, Bv, lm1, lmat0);
  const V lC = mt_adj(mat0, C, lm2, lmat0);
  const V lFr = lA, lFst0 = lA, lFsr = lA;
  const V lMbr = tsub(lB, lC), lMtr = tsub(lB, lC);
  const V lMsr = tadd(tmul(lB, eq1 * rbpNorm), tmul(lC, eq2 * rbpNorm));
  const V lMst = lMsr;
  const T lrbpNorm = (tdot(Msr, lB) + tdot(Mst, lB)) * eq1 + (tdot(Msr, lC) + tdot(Mst, lC)) * eq2;
  // twist
  const V lmt = mt_adj(mat, mt, lMtr, lmat);
  const T lpasi = kt * lmt.z;
  // bending
  Q lr = {0.0, vz()};
  T lks = 0.0, ltemp0 = 0.0, ltemp2 = 0.0, ltemp3 = 0.0, ltemp4 = 0.0;
  if (bend) {
    const V lmb = mt_adj(mat, mb, lMbr, lmat);
    const V lfs = mt_adj(mat, fs, lFsr, lmat);
    const V lms = mt_adj(mat, ms, lMsr, lmat);
    const T lsita = -0.5 * kb * s_fai * lmb.x + 0.5 * kb * c_fai * lmb.y -
                    0.5 * ks * c_fai * r0 * lfs.x - 0.5 * ks * s_fai * r0 * lfs.y +
                    0.5 * ks * s_fai * r0 * lms.x - 0.5 * ks * c_fai * r0 * lms.y;
    const T ls_fai = -0.5 * kb * sita * lmb.x - 0.5 * ks * sita * r0 * lfs.y +
                     0.5 * ks * sita * r0 * lms.x;
    const T lc_fai = 0.5 * kb * sita * lmb.y - 0.5 * ks * sita * r0 * lfs.x -
                     0.5 * ks * sita * r0 * lms.y;
    lks += -0.5 * c_fai * sita * r0 * lfs.x - 0.5 * s_fai * sita * r0 * lfs.y +
           0.5 * s_fai * sita * r0 * lms.x - 0.5 * c_fai * sita * r0 * lms.y;
    if (!(temp1 >= 1.0 || temp1 <= -1.0)) {
      const T ltemp1 = lsita * dacos(temp1);
      ltemp0 += ltemp1;
      ltemp2 -= ltemp1;
    }
    const T lsD = -(lc_fai * c_fai + ls_fai * s_fai) / sD;
    if (temp0 == 0.0) {
      lr.v.y += lc_fai / sD;
      lr.v.x -= ls_fai / sD;
      ltemp2 += lsD * dsqrt_half(sD);
    } else {
      ltemp3 += lc_fai / sD;
      ltemp4 += ls_fai / sD;
      const T lP = lsD * dsqrt_half(sD);
      ltemp0 += lP * temp2;
      ltemp2 += lP * temp0;
    }
  }
  lr.v.x += ltemp3 * r.v.z - ltemp4 * r.s + 2.0 * r.v.x * ltemp2;
  lr.v.z += ltemp3 * r.v.x + ltemp4 * r.v.y;
  lr.s += ltemp3 * r.v.y - ltemp4 * r.v.x;
  lr.v.y += ltemp3 * r.s + ltemp4 * r.v.z + 2.0 * r.v.y * ltemp2;
  if (temp0 != 0.0) {
    const T lc_p = 2.0 * dang * lpasi;
    lr.s += lc_p / sq0;
    const T lsq0 = -lc_p * c_p / sq0;
    ltemp0 += lsq0 * dsqrt_half(sq0);
  }
  lr.s += 2.0 * r.s * ltemp0;
  lr.v.z += 2.0 * r.v.z * ltemp0;
  // r = u qm, u = qm^-1 rp, rp = q1^-1 q2
  Q lu = {0.0, vz()}, lqm = {0.0, vz()}, liqm = {0.0, vz()}, lrp = {0.0, vz()}, liq1 = {0.0, vz()};
  Q lq2 = {0.0, vz()};
  qmul_adj(u, qm, lr, lu, lqm);
  qmul_adj(iqm, rp, lu, liqm, lrp);
  qmul_adj(iq1, q2, lrp, liq1, lq2);
  lqm.s += liqm.s;
  lqm.v = tsub(lqm.v, liqm.v);
  const Q lqm_m = qmat_adj(qm, lmat);
  lqm.s += lqm_m.s;
  lqm.v = tadd(lqm.v, lqm_m.v);
  // qm = (m0, m1, m2, 0) from rb
  V lrb = vz();
  T lN = 0.0;
  T lm1a = 0.0;
  const T lqm0 = lqm.s, lqm1 = lqm.v.x, lqm2 = lqm.v.y;
  if (xy2 != 0.0) {
    lm1a += lqm1 * rb.y / dS - lqm2 * rb.x / dS;
    lrb.y += lqm1 * m1a / dS;
    lrb.x += lqm2 * (-m1a) / dS;
    const T ldS = -(lqm1 * m1 + lqm2 * m2) / dS;
    const T lxy2 = ldS * dsqrt_half(dS);
    lrb.x += 2.0 * rb.x * lxy2;
    lrb.y += 2.0 * rb.y * lxy2;
  } else {
    lm1a += lqm1 - lqm2;
  }
  const T la0 = lqm0 * HALF_SQRT_2 * dsqrt_half(sa0);
  const T la1 = lm1a * (-HALF_SQRT_2) * dsqrt_half(sa1);
  lN += la0 * (1.0 - a0) / rbNorm + la1 * (1.0 - a1) / rbNorm;
  lrb.z += la0 / rbNorm - la1 / rbNorm;
  // Mst = rb_0 (-|Fst|) / |rb_0|
  V lFst = lFst0, lrb_0 = vz(), lrb_b0 = vz();
  if (R0 != 0.0) {
    lrb_0 = tadd(lrb_0, tmul(lMst, -FstN / R0));
    const T lFstN = -tdot(lMst, rb_0) / R0;
    const T lR0 = -tdot(lMst, Mst) / R0;
    lrb_0 = tadd(lrb_0, tmul(rb_0, lR0 / R0));
    if (FstN > 0.0) lFst = tadd(lFst, tmul(Fst, lFstN / FstN));
  }
  // Fst = rb_b0 (ks r0 gama) / |rb_b0|
  T lgama = 0.0;
  if (B != 0.0) {
    lrb_b0 = tadd(lrb_b0, tmul(lFst, gmul / B));
    const T lgmul = tdot(lFst, rb_b0) / B;
    const T lBn = -tdot(lFst, Fst) / B;
    lrb_b0 = tadd(lrb_b0, tmul(rb_b0, lBn / B));
    lks += lgmul * r0 * gama;
    lgama = lgmul * ks * r0;
  }
  // rb_b0 = rb x rb_0, rb_0 = rb x D0
  lrb = tadd(lrb, tcross(rb_0, lrb_b0));
  lrb_0 = tadd(lrb_0, tcross(lrb_b0, rb));
  lrb = tadd(lrb, tcross(D0, lrb_0));
  // gama = acos(rb . D0 / (|rb| |D0|))
  if (g_in) {
    const T lgc = lgama * dacos(gama_cos);
    lrb = tadd(lrb, tmul(D0, lgc / gden));
    lN += -lgc * gama_cos / gden * D0n;
  }
  // Fr = kr (rb (|rb| - r0)) / |rb|
  const V ldr = tmul(lFr, kr);
  g_kr = tdot(lFr, delta_r);
  lrb = tadd(lrb, tmul(ldr, sdr / rbNorm));
  lN += tdot(ldr, rb) / rbNorm - tdot(ldr, delta_r) / rbNorm;
  lrb = tadd(lrb, tmul(rb, lN / rbNorm));
  // rb = mat0 rbp
  V lrbp = m_adj(mat0, rbp, lrb, lmat0);
  if (rbpNorm > 0.0) lrbp = tadd(lrbp, tmul(rbp, lrbpNorm / rbpNorm));
  const Q lq1m = qmat_adj(q1, lmat0);
  g_q1 = tquat<double>{lq1m.s + liq1.s, tsub(lq1m.v, liq1.v)};
  g_q2 = lq2;
  g_pos2 = lrbp;
  g_pos1 = tneg(lrbp);
  g_ks = lks;
}

// Reverse mode of tfriction_law for a touching pair (the caller handles pairs
// apart: their outputs are constants): given the adjoints of (f1, m1, m2) and
// of the history written (lh), the adjoints of (x1, x2, v1, v2, w1, w2, the
// history read, k, ks, mu) on the slip branch the forward evaluation takes
// (mu_d = mu_s = mu).
EDL_HD void tfriction_law_vjp(tv3<double> pos1, tv3<double> vel1, tv3<double> w1, double r1,
                              tv3<double> pos2, tv3<double> vel2, tv3<double> w2, double r2,
                              double k, double ks, double mu, double dt, tv3<double> ffric,
                              bool slipping, tv3<double> lf1, tv3<double> lm1, tv3<double> lm2,
                              tv3<double> lh, tv3<double>* g)
{
  using namespace edl_rev;
  typedef double T;
  (void)slipping;  // mu_d = mu_s: the slip flag of the last step selects the same mu
  // ---- forward (tfriction_law's expressions) ----
  const V D = tsub(pos1, pos2);
  const T dist2 = tdot(D, D);
  const double eq = r1 + r2;
  const T dist = sqrt(dist2);
  const T sc = k * (dist - eq) / dist;
  const V force = tmul(D, sc);
  const double cr = r2 / eq;
  const V pos = tadd(pos2, tmul(D, cr));
  const V a1 = tsub(pos, pos1), a2 = tsub(pos, pos2);
  const V vp1 = tadd(vel1, tcross(w1, a1));
  const V vp2 = tadd(vel2, tcross(w2, a2));
  const V ds0 = tmul(tsub(vp2, vp1), dt);
  const T p = tdot(ds0, D), q = tdot(D, D);
  const T t = p / q;
  const V ds = tsub(ds0, tmul(D, t));
  const V ffr1 = tadd(ffric, tmul(ds, ks));
  const T Fn = tnorm(ffr1);
  const T Nn = tnorm(force);
  const bool slip = Fn > Nn * mu;
  T sc2 = 1.0;
  V ffr2 = ffr1;
  if (slip) {
    sc2 = (mu * Nn) / Fn;
    ffr2 = tmul(ffr1, sc2);
  }
  // ---- reverse ----
  // m1 = a1 x ffr2, m2 = a2 x (-ffr2), f1 = -force + ffr2, history out = ffr2
  V lffr2 = tadd(tadd(lh, lf1), tsub(tcross(lm1, a1), tcross(lm2, a2)));
  V la1 = tcross(ffr2, lm1);
  V la2 = tcross(tneg(ffr2), lm2);
  V lforce = tneg(lf1);
  V lffr1 = lffr2;
  T lmu = 0.0;
  if (slip) {
    lffr1 = tmul(lffr2, sc2);
    const T lsc2 = tdot(lffr2, ffr1);
    lmu += lsc2 * Nn / Fn;
    const T lNn = lsc2 * mu / Fn;
    const T lFn = -lsc2 * sc2 / Fn;
    if (Fn > 0.0) lffr1 = tadd(lffr1, tmul(ffr1, lFn / Fn));
    if (Nn > 0.0) lforce = tadd(lforce, tmul(force, lNn / Nn));
  }
  // ffr1 = ffric + ds ks
  const V lffric = lffr1;
  const V lds = tmul(lffr1, ks);
  const T lks = tdot(lffr1, ds);
  // ds = ds0 - D (ds0 . D) / (D . D)
  V lD = tmul(lds, -t);
  const T lt = -tdot(lds, D);
  const T lp = lt / q, lq = -lt * t / q;
  const V lds0 = tadd(lds, tmul(D, lp));
  lD = tadd(lD, tadd(tmul(ds0, lp), tmul(D, 2.0 * lq)));
  // ds0 = (vp2 - vp1) dt; vp = v + w x a
  const V lvp2 = tmul(lds0, dt), lvp1 = tmul(lds0, -dt);
  const V lw1 = tcross(a1, lvp1), lw2 = tcross(a2, lvp2);
  la1 = tadd(la1, tcross(lvp1, w1));
  la2 = tadd(la2, tcross(lvp2, w2));
  // a1 = pos - x1, a2 = pos - x2, pos = x2 + D r2 / eq
  const V lpos = tadd(la1, la2);
  V lx1 = tneg(la1), lx2 = tsub(lpos, la2);
  lD = tadd(lD, tmul(lpos, cr));
  // force = D k (|D| - eq) / |D|
  lD = tadd(lD, tmul(lforce, sc));
  const T lsc = tdot(lforce, D);
  const T lk = lsc * (dist - eq) / dist;
  const T ldist = lsc * (k - sc) / dist;
  lD = tadd(lD, tmul(D, 2.0 * ldist * dsqrt_half(dist)));
  // D = x1 - x2
  lx1 = tadd(lx1, lD);
  lx2 = tsub(lx2, lD);
  g[0] = lx1;
  g[1] = lx2;
  g[2] = lvp1;  // v1
  g[3] = lvp2;  // v2
  g[4] = lw1;
  g[5] = lw2;
  g[6] = lffric;
  g[7] = tv3<double>{lk, lks, lmu};
}
