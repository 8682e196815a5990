// esp_math.h - FP64 vector / quaternion arithmetic and the interaction laws,
// usable on host and device.
//
// Every expression keeps the association order of the reference's Foundation
// operators (Foundation/vec3.hpp:80-244, Matrix3.hpp:31-140,
// Quaternion.hpp:162-309) and interaction laws, and the library is compiled
// with -ffp-contract=off, so no multiply-add is fused: given equal inputs a law
// here rounds exactly like the reference's x86-64 -O2 build, except for acos,
// which is the device math library's (within 1 ulp of glibc's).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define ESP_HD __host__ __device__ __forceinline__

struct dv3 {
  double x, y, z;
};
struct dquat {
  double s;
  dv3 v;
};
struct dmat3 {
  double m[3][3];
};

ESP_HD dv3 mk3(double x, double y, double z) { return dv3{x, y, z}; }
ESP_HD dv3 add3(dv3 a, dv3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
ESP_HD dv3 sub3(dv3 a, dv3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
ESP_HD dv3 neg3(dv3 a) { return mk3(-a.x, -a.y, -a.z); }
ESP_HD dv3 mul3(dv3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
ESP_HD dv3 div3(dv3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
ESP_HD double dot3(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ESP_HD dv3 cross3(dv3 a, dv3 b)
{
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
ESP_HD double norm3(dv3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

ESP_HD dv3 mmul3(const dmat3& m, dv3 v)
{
  return mk3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
             m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
             m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}
// transpose(m) * v without materialising the transpose
ESP_HD dv3 mtmul3(const dmat3& m, dv3 v)
{
  return mk3(m.m[0][0] * v.x + m.m[1][0] * v.y + m.m[2][0] * v.z,
             m.m[0][1] * v.x + m.m[1][1] * v.y + m.m[2][1] * v.z,
             m.m[0][2] * v.x + m.m[1][2] * v.y + m.m[2][2] * v.z);
}
ESP_HD dquat qmul4(dquat a, dquat b)
{
  dquat r;
  r.s = a.s * b.s - dot3(a.v, b.v);
  r.v = add3(add3(mul3(b.v, a.s), mul3(a.v, b.s)), cross3(a.v, b.v));
  return r;
}
ESP_HD dquat qinv4(dquat a) { return dquat{a.s, neg3(a.v)}; }
ESP_HD dmat3 qmat4(dquat q)
{
  dmat3 r;
  const double s = q.s, x = q.v.x, y = q.v.y, z = q.v.z;
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
ESP_HD double esp_calc_angle(double s_in, double c_os)
{
  if (s_in > 0.0) return (c_os >= 1.0 || c_os <= -1.0) ? 0.0 : acos(c_os);
  if (s_in == 0.0) return 0.0;
  return (c_os >= 1.0 || c_os <= -1.0) ? 0.0 : -acos(c_os);
}

struct BondResult {
  dv3 f1;  // force on the lower-id particle; the other gets -f1
  dv3 m1, m2;
  double nForce, shForce, tMoment, bMoment;
};

// ARotBondedInteraction::calcForces (ARotBondedInteraction.cpp:307-448).
// D0 is the initial centre difference p2 - p1 fixed at bond creation.
ESP_HD void esp_rot_bond_law(dv3 pos1, dquat q1, double r1, dv3 pos2, dquat q2, double r2,
                             double kr, double ks, double kt, double kb, dv3 D0, BondResult& o)
{
  const double HALF_SQRT_2 = 0.5 * sqrt(2.0);
  double s_fai = 0.0, c_fai = 0.0;
  double sita = 0.0, pasi = 0.0;

  const dmat3 mat0 = qmat4(q1);
  const dv3 rbp = sub3(pos2, pos1);
  const double rbpNorm = norm3(rbp);
  const dv3 rb = mmul3(mat0, rbp);
  const double rbNorm = norm3(rb);
  const double r_0Norm = r1 + r2;
  const dv3 delta_r = div3(mul3(rb, rbNorm - r_0Norm), rbNorm);
  const dv3 Fr = mul3(delta_r, kr);

  const double gama_cos = dot3(rb, D0) / (rbNorm * norm3(D0));
  double gama = 0.0;
  if (gama_cos < 1.0 && gama_cos > -1.0) gama = acos(gama_cos);

  const dv3 rb_0 = cross3(rb, D0);
  const dv3 rb_b0 = cross3(rb, rb_0);
  const double rb_b0Norm = norm3(rb_b0);
  dv3 Fst = mk3(0.0, 0.0, 0.0);
  if (rb_b0Norm != 0) Fst = div3(mul3(rb_b0, ks * r_0Norm * gama), rb_b0Norm);

  dv3 Mst = mk3(0.0, 0.0, 0.0);
  const double rb_0Norm = norm3(rb_0);
  if (rb_0Norm != 0) Mst = div3(mul3(rb_0, -norm3(Fst)), rb_0Norm);

  const double m0 = HALF_SQRT_2 * sqrt((rbNorm + rb.z) / rbNorm);
  double m1 = -HALF_SQRT_2 * sqrt((rbNorm - rb.z) / rbNorm);
  double m2 = -m1;
  const double m3 = 0.0;
  const double rbX2PlusRbY2 = rb.x * rb.x + rb.y * rb.y;
  if (rbX2PlusRbY2 != 0.0) {
    const double denomSqrt = sqrt(rbX2PlusRbY2);
    m1 = m1 * rb.y / denomSqrt;
    m2 = m2 * rb.x / denomSqrt;
  }

  const dquat qm = {m0, {m1, m2, m3}};
  const dmat3 mat = qmat4(qm);  // applied transposed below (matTrans)
  const dquat rp = qmul4(qinv4(q1), q2);
  const dquat r = qmul4(qmul4(qinv4(qm), rp), qm);
  const double temp0 = r.s * r.s + r.v.z * r.v.z;
  if (temp0 == 0.0) {
    pasi = 0.0;
  } else {
    const double sqrtTemp0 = sqrt(temp0);
    const double c_pasi2 = r.s / sqrtTemp0;
    const double s_pasi2 = r.v.z / sqrtTemp0;
    pasi = 2.0 * esp_calc_angle(s_pasi2, c_pasi2);
  }
  const dv3 Mtr = mtmul3(mat, mk3(0.0, 0.0, kt * pasi));
  const double temp2 = r.v.x * r.v.x + r.v.y * r.v.y;
  const double temp3 = r.v.x * r.v.z + r.s * r.v.y;
  const double temp4 = r.v.y * r.v.z - r.s * r.v.x;

  dv3 Mbr = mk3(0, 0, 0), Fsr = mk3(0, 0, 0), Msr = mk3(0, 0, 0);
  if (temp2 != 0.0) {
    const double temp1 = temp0 - temp2;
    if (temp1 >= 1.0 || temp1 <= -1.0) sita = 0.0;
    else sita = acos(temp1);
    if (temp0 == 0.0) {
      const double sqrtDenom = sqrt(temp2);
      c_fai = r.v.y / sqrtDenom;
      s_fai = -r.v.x / sqrtDenom;
    } else {
      const double sqrtDenom = sqrt(temp0 * temp2);
      c_fai = temp3 / sqrtDenom;
      s_fai = temp4 / sqrtDenom;
    }
    Mbr = mtmul3(mat, mk3(-0.5 * kb * sita * s_fai, 0.5 * kb * sita * c_fai, 0.0));
    Fsr = mtmul3(mat, mk3(-0.5 * ks * c_fai * sita * r_0Norm, -0.5 * ks * s_fai * sita * r_0Norm,
                          0.0));
    Msr = mtmul3(mat, mk3(0.5 * ks * s_fai * sita * r_0Norm, -0.5 * ks * c_fai * sita * r_0Norm,
                          0.0));
  }

  const double eq_rad1 = r1 / (r1 + r2);
  const double eq_rad2 = r2 / (r1 + r2);
  o.f1 = mtmul3(mat0, add3(add3(Fr, Fst), Fsr));
  o.m1 = mtmul3(mat0, add3(add3(add3(Mbr, Mtr), mul3(mul3(Msr, eq_rad1), rbpNorm)),
                           mul3(mul3(Mst, eq_rad1), rbpNorm)));
  o.m2 = mtmul3(mat0, sub3(sub3(add3(mul3(mul3(Msr, eq_rad2), rbpNorm),
                                     mul3(mul3(Mst, eq_rad2), rbpNorm)),
                                Mbr),
                           Mtr));
  o.shForce = norm3(add3(Fst, Fsr));
  o.tMoment = norm3(Mtr);
  o.bMoment = norm3(Mbr);
  const dv3 D = sub3(pos1, pos2);
  const double dist = sqrt(dot3(D, D));
  o.nForce = ((dist < r_0Norm) ? -1.0 : 1.0) * norm3(Fr);
}

// CDamping::calcForces / CRotDamping::calcForces (Damping.hpp:87-107,
// RotDamping.hpp:58-84): the damped relative velocity solves
// v_rel = v - v_ref + s dt (f - v_rel visc m) by fixed-point iteration from
// zero while the squared norm of the last change, squared, exceeds 1e-12, at
// most maxit times; a run that reaches maxit gives v_rel = 0.  The reference
// velocity is zero (DampingPrms), so v - v_ref is v.  For the rotational form
// v, f, s, m are the angular velocity, moment, 1/I and I.
ESP_HD dv3 esp_damping_vrel(dv3 v, dv3 f, double s, double m, double visc, double dt, int maxit)
{
  dv3 v_rel = mk3(0.0, 0.0, 0.0);
  double error = 1.0;
  int count = 0;
  while ((error * error > 1e-12) & (count < maxit)) {
    count++;
    const dv3 v_old = v_rel;
    v_rel = add3(v, mul3(sub3(f, mul3(mul3(v_rel, visc), m)), s * dt));
    error = dot3(sub3(v_rel, v_old), sub3(v_rel, v_old));
  }
  if (count >= maxit) v_rel = mk3(0.0, 0.0, 0.0);
  return v_rel;
}

// CLocalDamping / CRotLocalDamping::calcForces (LocalDamping.hpp:82-117,
// RotLocalDamping.hpp:82-117): each component of the force (moment) against
// the sign of the velocity (angular velocity) component, zero where it is zero
ESP_HD double esp_local_damping_c(double f, double v)
{
  double d = fabs(f);
  if (v < 0) d *= -1.;
  if (v == 0) d = 0.0;
  return d;
}

// CRotBondedInteraction::broken (RotBondedInteraction.cpp:175-219), evaluated
// on the values calcForces cached for the same bond.
ESP_HD bool esp_bond_broken(double nForce, double shForce, double tMoment, double bMoment,
                            double max_n, double max_s, double max_t, double max_b,
                            double truncated, double beta1, double beta2)
{
  const double shearStress = shForce / max_s + beta2 * tMoment / max_t;
  const double normalStress = nForce / max_n + beta1 * bMoment / max_b;
  return (normalStress > truncated) || (shearStress > (1.0 - normalStress));
}

// Scaled constants of a rotational bond (RotBondedInteraction.cpp:122-161).
// maxes = {max_n, max_s, max_t, max_b} before size scaling.
struct BondConst {
  double kr, ks, kt, kb, max_n, max_s, max_t, max_b;
  dv3 D;
};
ESP_HD BondConst esp_bond_constants(dv3 pos1, double r1, dv3 pos2, double r2, double kr,
                                    double ks, double kt, double kb, const double* maxes,
                                    int scaling, int meanR_scaling, int do2d)
{
  BondConst c;
  c.D = sub3(pos2, pos1);
  const double dist = sqrt(dot3(c.D, c.D));
  double effR = 1.0, effL = 1.0, effA = 1.0, momentJ = 1.0, momentI = 1.0;
  if (scaling) {
    if (!do2d) {
      effR = meanR_scaling ? 0.5 * dist : fmin(r1, r2);
      effL = dist;
      effA = effR * effR;
      momentJ = 0.5 * effR * effR * effR * effR;
      momentI = 0.5 * momentJ;
    }
  }
  c.kr = kr * effA / effL;
  c.ks = ks * effA / effL;
  c.kb = kb * momentI / effL;
  c.kt = kt * momentJ / effL;
  c.max_n = maxes[0] * effA;
  c.max_s = maxes[1] * effA;
  c.max_b = maxes[3] * momentI / effR;
  c.max_t = maxes[2] * momentJ / effR;
  return c;
}

// Constructor scalings of the pair laws.
// CElasticInteraction (ElasticInteraction.cpp:25-46)
ESP_HD double esp_elastic_k(double k, int scaling, double r1, double r2, int do2d)
{
  double effR = 1.0, effA = 1.0, effL = 1.0;
  const double r0 = r1 + r2;
  if (scaling) {
    if (!do2d) effR = 0.5 * r0;
    effL = r0;
    effA = effR * effR;
  }
  return k * effA / effL;
}
// CRotElasticInteraction (RotElasticInteraction.cpp:40-57)
ESP_HD double esp_rotelastic_k(double kr, int scaling, double r1, double r2, int do2d)
{
  double f = 1.0;
  if (scaling && !do2d) f = 0.5 * (r1 + r2);
  return f * kr;
}
// CRotFrictionInteraction (RotFricInteraction.cpp:103-146): returns k, ks
ESP_HD void esp_friction_k(double k, double ks, int scaling, int meanR, double r1, double r2,
                           int do2d, double& ko, double& kso)
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

// CRotFrictionInteraction::calcSimpleForces (RotFricInteraction.cpp:171-231).
// ffric / slipping are the interaction's history (updated in place).  Returns
// touching; f1 is the total force on the lower-id particle (-Fn then +Ffric).
ESP_HD bool esp_rot_friction_law(dv3 pos1, dv3 vel1, dv3 w1, double r1, dv3 pos2, dv3 vel2,
                                 dv3 w2, double r2, double k, double ks, double mu_d,
                                 double mu_s, double dt, dv3& ffric, bool& slipping, dv3& f1,
                                 dv3& m1, dv3& m2)
{
  const double mu_current = slipping ? mu_d : mu_s;
  const dv3 D = sub3(pos1, pos2);
  double dist = dot3(D, D);
  const double eq_dist = r1 + r2;
  if (dist < (eq_dist * eq_dist)) {
    dist = sqrt(dist);
    const dv3 force = mul3(D, k * (dist - eq_dist) / dist);
    const dv3 pos = add3(pos2, mul3(D, r2 / eq_dist));
    const dv3 vp1 = add3(vel1, cross3(w1, sub3(pos, pos1)));
    const dv3 vp2 = add3(vel2, cross3(w2, sub3(pos, pos2)));
    dv3 ds = mul3(sub3(vp2, vp1), dt);
    ds = sub3(ds, mul3(D, dot3(ds, D) / dot3(D, D)));
    ffric = add3(ffric, mul3(ds, ks));
    const double FfricNorm = norm3(ffric);
    const double forceNorm = norm3(force);
    if (FfricNorm > forceNorm * mu_current) {
      ffric = mul3(ffric, (mu_d * forceNorm) / FfricNorm);
      slipping = true;
    } else {
      slipping = false;
    }
    m1 = cross3(sub3(pos, pos1), ffric);
    m2 = cross3(sub3(pos, pos2), neg3(ffric));
    f1 = add3(mul3(force, -1.0), ffric);
    return true;
  }
  ffric = mk3(0.0, 0.0, 0.0);
  slipping = false;
  return false;
}

// CRotFrictionInteraction::calcRigidBodyForces (RotFricInteraction.cpp:237-343),
// the rigid=True form.  The slip increment is the tangential part of the
// translational velocity difference plus the difference of the rotational
// velocities at the contact point; a sticking pair that touched in its last
// step (touching) first turns its stored shear force with the pair's rigid-body
// rotation (spin about the axis plus the axis' own rotation).  A slipping pair
// keeps slipping while the increment exceeds 1e-8 and otherwise sticks with
// the force it had.  Outputs as esp_rot_friction_law.
ESP_HD bool esp_rot_friction_rigid_law(dv3 pos1, dv3 vel1, dv3 w1, double r1, dv3 pos2, dv3 vel2,
                                       dv3 w2, double r2, double k, double ks, double mu_d,
                                       double mu_s, double dt, bool touching, dv3& ffric,
                                       bool& slipping, dv3& f1, dv3& m1, dv3& m2)
{
  const dv3 D = sub3(pos1, pos2);
  double dist = dot3(D, D);
  const double eq_dist = r1 + r2;
  if (dist < (eq_dist * eq_dist)) {
    dist = sqrt(dist);
    const dv3 force = mul3(D, k * (dist - eq_dist) / dist);
    const dv3 pos = add3(pos2, mul3(D, r2 / eq_dist));
    const dv3 vp1_rot = cross3(w1, sub3(pos, pos1));
    const dv3 vp2_rot = cross3(w2, sub3(pos, pos2));
    const dv3 dv_trs = sub3(vel2, vel1);
    const dv3 dv_trs_s = sub3(dv_trs, mul3(D, dot3(dv_trs, D) / dot3(D, D)));
    const dv3 ds = mul3(add3(sub3(vp2_rot, vp1_rot), dv_trs_s), dt);
    const double forceNorm = norm3(force);
    if (!slipping) {
      if (!touching) {
        ffric = mk3(0.0, 0.0, 0.0);
      } else {
        const dv3 rbp = sub3(pos2, pos1);
        const dv3 vbp = sub3(vel2, vel1);
        const double rbp0 = norm3(rbp);
        const dv3 om_s = mul3(add3(w1, w2), 0.5);
        const dv3 om_spin = div3(mul3(rbp, dot3(om_s, rbp)), rbp0 * rbp0);
        const dv3 om_m = div3(cross3(rbp, vbp), rbp0 * rbp0);
        ffric = add3(ffric, mul3(cross3(add3(om_spin, om_m), ffric), dt));
      }
      if (norm3(add3(ffric, mul3(ds, ks))) > forceNorm * mu_s) {
        slipping = true;
        ffric = div3(mul3(ds, mu_d * forceNorm), norm3(ds));
      } else {
        ffric = add3(ffric, mul3(ds, ks));
      }
    } else if (norm3(ds) > 1.0e-8) {
      ffric = div3(mul3(ds, mu_d * forceNorm), norm3(ds));
    } else {
      slipping = false;
    }
    m1 = cross3(sub3(pos, pos1), ffric);
    m2 = cross3(sub3(pos, pos2), neg3(ffric));
    f1 = add3(mul3(force, -1.0), ffric);
    return true;
  }
  ffric = mk3(0.0, 0.0, 0.0);
  slipping = false;
  return false;
}

// 1 / particle volume (CParticle::m_div_vol, Particle.cpp:52)
ESP_HD double esp_div_vol(double r) { return 1.0 / (4.0 / 3.0 * M_PI * r * r * r); }

// BrittleBeamSCInteraction::broken (BrittleBeamSC.cpp:103-148): truncated
// Mohr-Coulomb test on the mean of the two particles' stresses
// (CParticle::sigma = m_sigma / volume, Particle.cpp:405-408; sig1/sig2 are the
// unscaled sums, row-major) in the frame of the current bond direction.
ESP_HD bool esp_sc_broken(dv3 pos1, double r1, const double* sig1, dv3 pos2, double r2,
                          const double* sig2, double sigma3, double cohesion, double tan_angle)
{
  const dv3 d = sub3(pos2, pos1);
  const dv3 n = div3(d, norm3(d));
  dv3 s;
  if (n.x != 1.0 || n.y != 0.0 || n.z != 0.0) {
    const dv3 c = cross3(n, mk3(1.0, 0.0, 0.0));
    s = div3(c, norm3(c));
  } else {
    const dv3 c = cross3(n, mk3(0.0, 1.0, 0.0));
    s = div3(c, norm3(c));
  }
  const dv3 t = cross3(n, s);
  // Matrix3(n,s,t).trans(): the axes as rows
  const double tm[3][3] = {{n.x, n.y, n.z}, {s.x, s.y, s.z}, {t.x, t.y, t.z}};
  const double dv1 = esp_div_vol(r1), dv2 = esp_div_vol(r2);
  double avg[3][3], sym[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) avg[i][j] = 0.5 * (sig1[3 * i + j] * dv1 + sig2[3 * i + j] * dv2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sym[i][j] = 0.5 * (avg[i][j] + avg[j][i]);
  // first row of tm . sym . tm^T (Matrix3::dot sums from zero, Matrix3.cpp:320-333)
  double a0[3];
  for (int k = 0; k < 3; k++) {
    double v = 0.0;
    for (int m = 0; m < 3; m++) v += tm[0][m] * sym[m][k];
    a0[k] = v;
  }
  double rot0[3];
  for (int j = 0; j < 3; j++) {
    double v = 0.0;
    for (int k = 0; k < 3; k++) v += a0[k] * tm[j][k];
    rot0[j] = v;
  }
  const double sigma_n = rot0[0];
  const double tau = sqrt(rot0[1] * rot0[1] + rot0[2] * rot0[2]);
  return (sigma_n < -1.0 * sigma3) || (tau > cohesion + sigma_n * tan_angle);
}
