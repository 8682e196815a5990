// esp_diff.hip - the differentiable, batched engine of include/esp_diff.h.
//
// Forward: the reference's time step for the calibration model class over one
// particle set holding any number of specimens (per-specimen parameters),
// with the main engine's neighbour-search kernels (cell binning, 27-cell pair
// search) and three step kernels:
//   k_bond_fwd    per bond: ARotBondedInteraction::calcForces + the break test
//                 (the bond leaves at the next step)
//   k_fric_fwd    per neighbour pair: RotFriction unless a live bond excludes
//                 it; shear history read from one buffer, written to the other
//   k_gather_fwd  per particle: walls, gravity, then its interactions in a
//                 fixed order (pairs by key, then bonds), the CParticle /
//                 CRotParticle update, the rebuild test
// plus a per-specimen, fixed-order sum of the measured wall's force at sample
// steps.  Checkpoints every K steps keep the state, the shear histories and
// the bond break steps; the neighbour lists ("epochs", one per rebuild) are
// immutable and kept for the whole run.
//
// Backward: segments in reverse order; each is recomputed forward from its
// checkpoint (recorded epochs, no search) storing the state of every step,
// then swept backwards:
//   k_adj_particle  the integrator's adjoint (symplectic Euler, quaternion
//                   step and normalisation), the walls' and the loss's terms
//   k_adj_bond      vector-Jacobian product of the bond law (16 inputs), by
//                   its hand-written reverse mode
//   k_adj_fric      vector-Jacobian product of the friction law (24 inputs,
//                   the history among them, on the recorded slip branch), by
//                   its hand-written reverse mode
//   k_adj_gather    per particle, the interactions' adjoints in a fixed order,
//                   and the parameter gradients credited to the first particle
// The laws' Jacobian columns come from their dual-number evaluation
// (esp_difflaw.h); every reduction has a fixed order (no atomics on doubles).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/esp_diff.h"
#include "esp_difflaw.h"
#include "esp_internal.h"
#include "esp_math.h"

namespace {

struct DiffError : std::runtime_error {
  int code;
  DiffError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DCHK(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw DiffError(ESP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
  void reserve(size_t n)
  {
    if (n <= cap && p) return;
    if (p) {
      DCHK(hipDeviceSynchronize());
      DCHK(hipFree(p));
    }
    p = nullptr;
    const size_t c = std::max<size_t>(n, 16);
    DCHK(hipMalloc(&p, c * sizeof(T)));
    cap = c;
  }
  void swap(DevBuf& o)
  {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

thread_local std::string g_diff_create_err;

__device__ __forceinline__ long long dpair_key(int a, int b)
{
  return (long long)(((unsigned long long)(unsigned)a << 32) | (unsigned long long)(unsigned)b);
}

// Branch digest of a step: each wave hashes (its first index, its lanes'
// branch bits) and adds the hash into the step's word; the sum is independent
// of the order the waves finish in.  Two runs with equal digests took the same
// contact, slip and wall-contact branches (pair and particle indices of the
// same neighbour-table epochs).
__device__ __forceinline__ unsigned long long dmix(unsigned long long z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ void wave_digest(unsigned long long* dig, unsigned long long tag,
                                            unsigned long long m0, unsigned long long m1)
{
  const int first = __ffsll((long long)__ballot(1)) - 1;
  if ((int)__lane_id() == first) atomicAdd(dig, dmix(dmix(tag) ^ m0) + dmix(m1 ^ ~tag));
}

// ---------------------------------------------------------------------------
// device views
// ---------------------------------------------------------------------------
struct DState {
  double4* x;  // position, radius
  double4* v;  // velocity, 1/mass
  double4* w;  // angular velocity, 1/inertia
  double4* q;  // orientation
  const int* spec;
  int n;
};

#define DIFF_MAX_WALLS 4
struct DWall {
  double o[3], n[3], k;
  int measure;
};
struct DStep {
  int nw;
  DWall wall[DIFF_MAX_WALLS];
  int has_g;
  double g[3];
  double dt;
  int step;
  const double* wo;  // wall origins per step [step][wall][3] (forward), or wall[].o
};

__device__ __forceinline__ dv3 wall_origin(const DStep& sp, int k, int step)
{
  if (sp.wo) {
    const double* o = sp.wo + ((size_t)step * sp.nw + k) * 3;
    return mk3(o[0], o[1], o[2]);
  }
  return mk3(sp.wall[k].o[0], sp.wall[k].o[1], sp.wall[k].o[2]);
}

// Device-driven stepping: the kernels of a forward step read the step they
// belong to from `step`; a step whose displacement test asks for a neighbour
// search records itself in `halt`, and every later step of the queue (or of a
// replayed graph) returns at once; steps past `end` return too.
struct DRun {
  int step, halt, end, pad;
};

__device__ __forceinline__ bool run_skip(const DRun* R, int& step)
{
  step = R->step;
  const int h = R->halt;
  return step > R->end || (h != 0 && h < step);
}

// Largest block of the per-step kernels.  The law kernels declare it as their
// launch bound, which gives them 256 VGPRs a lane (the default bound of 1024
// threads caps them at 128, and the bond law then spills to scratch memory).
#define DIFF_MAX_BS 512

__global__ void k_set_run(DRun* R, int step, int halt, int end)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    R->step = step;
    R->halt = halt;
    R->end = end;
  }
}

// the last launch of a step
__global__ void k_advance(DRun* R)
{
  int step;
  if (blockIdx.x != 0 || threadIdx.x != 0 || run_skip(R, step)) return;
  R->step = step + 1;
}

struct DBonds {
  int nb;
  const int2* ij;
  const double4* D0;   // initial p2 - p1, |D0|
  const double4* rel;  // relative break values (n, s, t, b)
  int* bstep;          // step whose criterion broke the bond (INT_MAX: intact)
  const double2* tkb;  // per specimen (kt, kb)
};

struct DCands {
  int nc;
  const int2* ij;
  const int* bond;  // bond between the two, or -1
};

__device__ __forceinline__ tv3<double> dx3(double4 a) { return tv3<double>{a.x, a.y, a.z}; }

// ---------------------------------------------------------------------------
// tangent-linear (forward-mode) kernels: the derivative of one forward run
// along a parameter direction, on the branch the run takes.  T holds the
// state's tangent (x, v, w, q), th the shear histories' tangents; the laws run
// on duals seeded with the state's tangents and the parameters' direction.
// ---------------------------------------------------------------------------
struct DTan {
  double4* x;
  double4* v;
  double4* w;
  double4* q;
};

__device__ __forceinline__ tv3<esp_dual> dd3(double4 a, double4 t)
{
  return tv3<esp_dual>{esp_dual(a.x, t.x), esp_dual(a.y, t.y), esp_dual(a.z, t.z)};
}
__device__ __forceinline__ tquat<esp_dual> ddq(double4 a, double4 t)
{
  return tquat<esp_dual>{esp_dual(a.x, t.x), {esp_dual(a.y, t.y), esp_dual(a.z, t.z), esp_dual(a.w, t.w)}};
}

__global__ __launch_bounds__(DIFF_MAX_BS) void k_bond_tan(DState S, DTan T, DBonds B, const double4* theta, const double4* dtheta,
                           const DRun* run, double* bt)
{
  int step;
  if (run_skip(run, step)) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  double* o = bt + 9 * (size_t)b;
  if (B.bstep[b] < step) {
    for (int k = 0; k < 9; k++) o[k] = 0.0;
    return;
  }
  const int2 e = B.ij[b];
  const double4 D = B.D0[b];
  const int s = S.spec[e.x];
  const double4 th = theta[s], dth = dtheta[s];
  const double2 tk = B.tkb[s];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y];
  esp_dual kr, ks;
  double kt, kb;
  tbond_k<esp_dual>(esp_dual(th.x, dth.x), esp_dual(th.y, dth.y), tk.x, tk.y, D.w, P1.w, P2.w, 1,
                    1, 0, kr, ks, kt, kb);
  TBondResult<esp_dual> r;
  tbond_law<esp_dual>(dd3(P1, T.x[e.x]), ddq(S.q[e.x], T.q[e.x]), P1.w, dd3(P2, T.x[e.y]),
                      ddq(S.q[e.y], T.q[e.y]), P2.w, kr, ks, esp_dual(kt), esp_dual(kb),
                      tv3<esp_dual>{esp_dual(D.x), esp_dual(D.y), esp_dual(D.z)}, r);
  o[0] = r.f1.x.d; o[1] = r.f1.y.d; o[2] = r.f1.z.d;
  o[3] = r.m1.x.d; o[4] = r.m1.y.d; o[5] = r.m1.z.d;
  o[6] = r.m2.x.d; o[7] = r.m2.y.d; o[8] = r.m2.z.d;
}

__global__ __launch_bounds__(DIFF_MAX_BS) void k_fric_tan(DState S, DTan T, DCands C, const int* bstep, const double4* theta,
                           const double4* dtheta, const double4* H0, const double4* H1,
                           double4* TH0, double4* TH1, const DRun* run, double dt, double* ct)
{
  int step;
  if (run_skip(run, step)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C.nc) return;
  // step s reads the histories of buffer s & 1 and writes the other
  const double4* hin = (step & 1) ? H1 : H0;
  const double4* thin = (step & 1) ? TH1 : TH0;
  double4* thout = (step & 1) ? TH0 : TH1;
  double* o = ct + 9 * (size_t)c;
  const int bb = C.bond[c];
  if (bb >= 0 && bstep[bb] >= step) {
    for (int k = 0; k < 9; k++) o[k] = 0.0;
    thout[c] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  const int2 e = C.ij[c];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y];
  const int s = S.spec[e.x];
  const double4 th = theta[s], dth = dtheta[s];
  esp_dual kk, kks;
  tfriction_k<esp_dual>(esp_dual(th.x, dth.x), esp_dual(th.y, dth.y), 1, 1, P1.w, P2.w, 0, kk, kks);
  const double4 h = hin[c], dh = thin[c];
  tv3<esp_dual> ffr{esp_dual(h.x, dh.x), esp_dual(h.y, dh.y), esp_dual(h.z, dh.z)}, f1, m1, m2;
  bool slipping = (((int)h.w) & 1) != 0;
  const esp_dual mu(th.w, dth.w);
  tfriction_law<esp_dual>(dd3(P1, T.x[e.x]), dd3(S.v[e.x], T.v[e.x]), dd3(S.w[e.x], T.w[e.x]), P1.w,
                          dd3(P2, T.x[e.y]), dd3(S.v[e.y], T.v[e.y]), dd3(S.w[e.y], T.w[e.y]), P2.w,
                          kk, kks, mu, mu, dt, ffr, slipping, f1, m1, m2);
  thout[c] = make_double4(ffr.x.d, ffr.y.d, ffr.z.d, 0.0);
  o[0] = f1.x.d; o[1] = f1.y.d; o[2] = f1.z.d;
  o[3] = m1.x.d; o[4] = m1.y.d; o[5] = m1.z.d;
  o[6] = m2.x.d; o[7] = m2.y.d; o[8] = m2.z.d;
}

// ---------------------------------------------------------------------------
// forward kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DIFF_MAX_BS) void k_bond_fwd(DState S, DBonds B, const double4* theta, const DRun* run, double* bout)
{
  int step;
  if (run_skip(run, step)) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  double* o = bout + 9 * (size_t)b;
  const int bs = B.bstep[b];
  // removed after the step whose criterion broke it: no record (the gathers
  // skip a bond whose break step is before theirs)
  if (bs < step) return;
  const int2 e = B.ij[b];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y], Q1 = S.q[e.x], Q2 = S.q[e.y];
  const double4 D = B.D0[b];
  const int s = S.spec[e.x];
  const double4 th = theta[s];
  const double2 tk = B.tkb[s];
  double kr, ks, kt, kb;
  tbond_k<double>(th.x, th.y, tk.x, tk.y, D.w, P1.w, P2.w, 1, 1, 0, kr, ks, kt, kb);
  TBondResult<double> r;
  tbond_law<double>(dx3(P1), tquat<double>{Q1.x, {Q1.y, Q1.z, Q1.w}}, P1.w, dx3(P2),
                    tquat<double>{Q2.x, {Q2.y, Q2.z, Q2.w}}, P2.w, kr, ks, kt, kb,
                    tv3<double>{D.x, D.y, D.z}, r);
  o[0] = r.f1.x; o[1] = r.f1.y; o[2] = r.f1.z;
  o[3] = r.m1.x; o[4] = r.m1.y; o[5] = r.m1.z;
  o[6] = r.m2.x; o[7] = r.m2.y; o[8] = r.m2.z;
  if (bs != INT_MAX) return;
  // CRotBondedInteraction::broken on the size-scaled break values
  // (RotBondedInteraction.cpp:144-157, 175-219): cohesion times the relative
  // values, scaled like esp_bond_constants
  const double4 rl = B.rel[b];
  const double effR = 0.5 * D.w, effA = effR * effR;
  const double momentJ = 0.5 * effR * effR * effR * effR, momentI = 0.5 * momentJ;
  const double mx_n = (th.z * rl.x) * effA, mx_s = (th.z * rl.y) * effA;
  const double mx_b = (th.z * rl.w) * momentI / effR, mx_t = (th.z * rl.z) * momentJ / effR;
  if (esp_bond_broken(r.nForce, r.shForce, r.tMoment, r.bMoment, mx_n, mx_s, mx_t, mx_b, 1.0, 1.0,
                      1.0))
    B.bstep[b] = step;
}

// History flag bits (w of a history): 1 slipping, 2 touching, 4 zero in both
// buffers (the pair has been apart for at least the last step, so neither
// buffer needs writing until it touches).  act[c]: the record of pair c holds
// this step's (nonzero) force; the gathers read only such records.
__global__ void k_fric_fwd(DState S, DCands C, const int* bstep, const double4* theta,
                           double4* H0, double4* H1, const DRun* run, double dt, double* cout,
                           unsigned char* act, unsigned long long* dig_base)
{
  int step;
  if (run_skip(run, step)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C.nc) return;
  // step s reads the histories of buffer s & 1 and writes the other
  double4* hin = (step & 1) ? H1 : H0;
  double4* hout = (step & 1) ? H0 : H1;
  unsigned long long* dig = dig_base ? dig_base + step : nullptr;
  const int bb = C.bond[c];
  if (bb >= 0 && bstep[bb] >= step) {
    // excluded while the bond lives: no record, and no history is read until
    // the step after its break, which reads the zero history written here at
    // the break step
    act[c] = 0;
    if (bstep[bb] == step) hout[c] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  const int2 e = C.ij[c];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y];
  const double4 h = hin[c];
  // the law's contact test, before the velocities are loaded
  const tv3<double> Dq = tsub(dx3(P1), dx3(P2));
  const double eq = P1.w + P2.w;
  const bool touching = tdot(Dq, Dq) < (eq * eq);
  tv3<double> ffr{h.x, h.y, h.z}, f1, m1, m2;
  bool slipping = false;
  if (touching) {
    const double4 V1 = S.v[e.x], V2 = S.v[e.y], W1 = S.w[e.x], W2 = S.w[e.y];
    const double4 th = theta[S.spec[e.x]];
    double kk, kks;
    tfriction_k<double>(th.x, th.y, 1, 1, P1.w, P2.w, 0, kk, kks);
    slipping = (((int)h.w) & 1) != 0;
    tfriction_law<double>(dx3(P1), dx3(V1), dx3(W1), P1.w, dx3(P2), dx3(V2), dx3(W2), P2.w, kk, kks,
                          th.w, th.w, dt, ffr, slipping, f1, m1, m2);
  }
  if (dig) {
    const unsigned long long mt = __ballot(touching), ms = __ballot(touching && slipping);
    wave_digest(dig, (unsigned long long)(c & ~63), mt, ms);
  }
  if (!touching) {
    // apart: history cleared, no force
    act[c] = 0;
    const int fw = (int)h.w;
    if ((fw & 3) != 0 || h.x != 0.0 || h.y != 0.0 || h.z != 0.0) {
      hout[c] = make_double4(0.0, 0.0, 0.0, 0.0);
    } else if (!(fw & 4)) {
      hin[c] = make_double4(0.0, 0.0, 0.0, 4.0);
      hout[c] = make_double4(0.0, 0.0, 0.0, 4.0);
    }
    return;
  }
  double* o = cout + 9 * (size_t)c;
  hout[c] = make_double4(ffr.x, ffr.y, ffr.z, (double)((slipping ? 1 : 0) | 2));
  o[0] = f1.x; o[1] = f1.y; o[2] = f1.z;
  o[3] = m1.x; o[4] = m1.y; o[5] = m1.z;
  o[6] = m2.x; o[7] = m2.y; o[8] = m2.z;
  act[c] = 1;
}

// per particle: the measured wall's normal force on it (sample steps), walls,
// gravity, interactions in CSR order, then the update (integrate = 1) or only
// the total force and moment (integrate = 0 writes fm)
__global__ void k_gather_fwd(DState S, const double4* mass_old, const int* inc_off,
                             const int* inc_ent, int ncand, const double* cout, const double* bout,
                             const unsigned char* cact, const int* bstep, DStep sp, int every,
                             double* wfn_s, DRun* run, unsigned long long thresh,
                             unsigned long long* dig_base, DTan T, const double* ct,
                             const double* bt, double* wft_s)
{
  int step;
  if (run_skip(run, step)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.n) return;
  const bool sample = every > 0 && step % every == 0;
  double* wfn = sample ? wfn_s : nullptr;
  double* wft = sample ? wft_s : nullptr;
  unsigned long long* dig = dig_base ? dig_base + step : nullptr;
  const double4 P = S.x[i];
  const dv3 x0 = mk3(P.x, P.y, P.z);
  const double r = P.w;
  dv3 F = mk3(0.0, 0.0, 0.0), M = mk3(0.0, 0.0, 0.0);
  dv3 tF = mk3(0.0, 0.0, 0.0), tM = mk3(0.0, 0.0, 0.0);  // tangents (T.x set)
  double wn = 0.0, twn = 0.0;
  for (int k = 0; k < sp.nw; k++) {  // CElasticWallInteraction::calcForces
    const DWall& W = sp.wall[k];
    const dv3 n = mk3(W.n[0], W.n[1], W.n[2]);
    const double dist = dot3(sub3(x0, wall_origin(sp, k, step)), n);
    if (dig) wave_digest(dig, ((unsigned long long)(i & ~63) << 3) | (unsigned long long)(k + 1),
                         __ballot(dist < r), 0ull);
    if (dist < r) {
      double f = 1.0;
      f *= 3.141592654 * r;
      const dv3 force = mul3(n, (f * W.k) * (r - dist));
      F = add3(F, force);
      if (W.measure) wn = dot3(force, n);
      if (T.x) {  // d force = -k' n (n . dx)
        const double4 Tx = T.x[i];
        const dv3 tf = mul3(n, -((f * W.k) * dot3(n, mk3(Tx.x, Tx.y, Tx.z))));
        tF = add3(tF, tf);
        if (W.measure) twn = dot3(tf, n);
      }
    }
  }
  if (wfn) wfn[i] = wn;
  if (wft) wft[i] = twn;
  const double4 Om = mass_old[i];
  if (sp.has_g) F = add3(F, mul3(mk3(sp.g[0], sp.g[1], sp.g[2]), Om.w));
  // The interactions in CSR order.  Up to GATHER_FWD_RUN entries of the run
  // load their record references together, then their activity together, then
  // the records of each chunk of GATHER_FWD_CH entries that holds an active one
  // (bonds, which are nearly all active, sit at the end of a particle's run);
  // the sums stay sequential.  A record with no force this step (a pair apart
  // or excluded by its bond, a bond broken in an earlier step) is skipped: it
  // would add zeros, and a sum that starts at +0 is not changed by a zero.
  constexpr int GATHER_FWD_RUN = 8, GATHER_FWD_CH = 4;
  const int e0 = inc_off[i], e1 = inc_off[i + 1];
  for (int r0 = e0; r0 < e1; r0 += GATHER_FWD_RUN) {
    int ent[GATHER_FWD_RUN];
#pragma unroll
    for (int k = 0; k < GATHER_FWD_RUN; k++) ent[k] = r0 + k < e1 ? inc_ent[r0 + k] : -1;
    unsigned act = 0;
#pragma unroll
    for (int k = 0; k < GATHER_FWD_RUN; k++) {
      const int rec = ent[k] >> 1;
      const bool on =
          ent[k] >= 0 && (rec < ncand ? cact[rec] != 0 : bstep[rec - ncand] >= step);
      act |= (on ? 1u : 0u) << k;
    }
#pragma unroll
    for (int c0 = 0; c0 < GATHER_FWD_RUN; c0 += GATHER_FWD_CH) {
      if (!((act >> c0) & ((1u << GATHER_FWD_CH) - 1u))) continue;
      dv3 f[GATHER_FWD_CH], m[GATHER_FWD_CH];
#pragma unroll
      for (int k = 0; k < GATHER_FWD_CH; k++) {
        if (!((act >> (c0 + k)) & 1u)) continue;
        const int rec = ent[c0 + k] >> 1, side = ent[c0 + k] & 1;
        const double* q = rec < ncand ? cout + 9 * (size_t)rec : bout + 9 * (size_t)(rec - ncand);
        f[k] = mk3(q[0], q[1], q[2]);
        m[k] = side == 0 ? mk3(q[3], q[4], q[5]) : mk3(q[6], q[7], q[8]);
      }
#pragma unroll
      for (int k = 0; k < GATHER_FWD_CH; k++) {
        if (!((act >> (c0 + k)) & 1u)) continue;
        if ((ent[c0 + k] & 1) == 0) F = add3(F, f[k]);
        else F = add3(F, mk3(-f[k].x, -f[k].y, -f[k].z));
        M = add3(M, m[k]);
      }
      if (T.x) {
#pragma unroll
        for (int k = 0; k < GATHER_FWD_CH; k++) {
          if (!((act >> (c0 + k)) & 1u)) continue;
          const int rec = ent[c0 + k] >> 1, side = ent[c0 + k] & 1;
          const double* u = rec < ncand ? ct + 9 * (size_t)rec : bt + 9 * (size_t)(rec - ncand);
          if (side == 0) {
            tF = add3(tF, mk3(u[0], u[1], u[2]));
            tM = add3(tM, mk3(u[3], u[4], u[5]));
          } else {
            tF = sub3(tF, mk3(u[0], u[1], u[2]));
            tM = add3(tM, mk3(u[6], u[7], u[8]));
          }
        }
      }
    }
  }
  // CRotParticle::integrate + rescale, CParticle::integrate (RotParticle.cpp:
  // 277-321, Particle.cpp:257-274), the expression order of k_gather_integrate
  const double dt = sp.dt;
  const double4 Vv = S.v[i], A = S.w[i], Q = S.q[i];
  dv3 v = mk3(Vv.x, Vv.y, Vv.z), x = x0;
  dv3 w = mk3(A.x, A.y, A.z);
  dquat qq = dquat{Q.x, mk3(Q.y, Q.z, Q.w)};
  w = add3(w, mul3(M, dt * A.w));
  const dquat dq = qmul4(dquat{0, w}, qq);
  const double h = dt / 2.0;
  qq.s += h * dq.s;
  qq.v = add3(qq.v, mul3(dq.v, h));
  v = add3(v, mul3(mul3(F, Vv.w), dt));
  x = add3(x, mul3(v, dt));
  const double module = sqrt(qq.s * qq.s + qq.v.x * qq.v.x + qq.v.y * qq.v.y + qq.v.z * qq.v.z);
  if (module != 0.0) {
    const double inverse = 1.0 / module;
    qq.s = inverse * qq.s;
    qq.v = mul3(qq.v, inverse);
  }
  if (T.x) {
    // tangents of w1 = w + M dt/I, qt = q + h (0, w1) q, q1 = qt/|qt|,
    // v1 = v + F dt/m, x1 = x + v1 dt (q0, w: the state before the step)
    const double4 Tx = T.x[i], Tv = T.v[i], Tw = T.w[i], Tq = T.q[i];
    const dquat q0 = dquat{Q.x, mk3(Q.y, Q.z, Q.w)}, tq0 = dquat{Tq.x, mk3(Tq.y, Tq.z, Tq.w)};
    const dv3 tw1 = add3(mk3(Tw.x, Tw.y, Tw.z), mul3(tM, dt * A.w));
    const dquat a1 = qmul4(dquat{0, tw1}, q0), a2 = qmul4(dquat{0, w}, tq0);
    const dquat tqt = dquat{tq0.s + h * (a1.s + a2.s), add3(tq0.v, mul3(add3(a1.v, a2.v), h))};
    dquat tq1 = tqt;
    if (module != 0.0) {
      const double pr = qq.s * tqt.s + dot3(qq.v, tqt.v);
      tq1 = dquat{(tqt.s - qq.s * pr) / module, div3(sub3(tqt.v, mul3(qq.v, pr)), module)};
    }
    const dv3 tv1 = add3(mk3(Tv.x, Tv.y, Tv.z), mul3(mul3(tF, Vv.w), dt));
    const dv3 tx1 = add3(mk3(Tx.x, Tx.y, Tx.z), mul3(tv1, dt));
    T.x[i] = make_double4(tx1.x, tx1.y, tx1.z, 0.0);
    T.v[i] = make_double4(tv1.x, tv1.y, tv1.z, 0.0);
    T.w[i] = make_double4(tw1.x, tw1.y, tw1.z, 0.0);
    T.q[i] = make_double4(tq1.s, tq1.v.x, tq1.v.y, tq1.v.z);
  }
  S.x[i] = make_double4(x.x, x.y, x.z, r);
  S.v[i] = make_double4(v.x, v.y, v.z, Vv.w);
  S.w[i] = make_double4(w.x, w.y, w.z, A.w);
  S.q[i] = make_double4(qq.s, qq.v.x, qq.v.y, qq.v.z);
  const dv3 d = sub3(x, mk3(Om.x, Om.y, Om.z));
  const unsigned long long m = (unsigned long long)__double_as_longlong(dot3(d, d));
  if (m > thresh) atomicMax(&run->halt, step);  // a neighbour search before the next step
}

// per specimen: fixed-order sum of its particles' measured wall forces, at the
// sample steps (the sum of sample k goes to out_base[k * nspec + s])
__global__ void k_spec_sum(const DRun* run, int every, const double* val, const int* spec_start,
                           int nspec, double* out_base)
{
  __shared__ double red[1024];
  int step;
  if (run_skip(run, step) || every <= 0 || step % every != 0) return;
  double* out = out_base + (size_t)(step / every - 1) * nspec;
  const int s = blockIdx.x;
  const int a = spec_start[s], b = spec_start[s + 1];
  double acc = 0.0;
  for (int i = a + (int)threadIdx.x; i < b; i += blockDim.x) acc = acc + val[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)blockDim.x; k++) t = t + red[k];
    out[s] = t;
  }
}

// ---------------------------------------------------------------------------
// neighbour lists of an epoch
// ---------------------------------------------------------------------------
__device__ __forceinline__ int dbsearch(const long long* a, int n, long long key)
{
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const long long v = a[mid];
    if (v == key) return mid;
    if (v < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

// candidate pairs -> pair slots, their bond and the pair of the previous epoch
// with the same key (whose shear history it takes over)
__global__ void k_epoch_pairs(int nc, const int* cp, const int* cq, const long long* ckey,
                              const long long* bkey, int nb, const long long* pkey, int np,
                              int2* ij, int* bond, int* src)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  ij[c] = make_int2(cp[c], cq[c]);
  bond[c] = dbsearch(bkey, nb, ckey[c]);
  src[c] = pkey ? dbsearch(pkey, np, ckey[c]) : -1;
}

__global__ void k_carry_hist(int nc, const int* src, const double4* hold, double4* hnew)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int s = src[c];
  hnew[c] = s >= 0 ? hold[s] : make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ void k_inc_count2(int nrec, const int2* ij, int* cnt)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrec) return;
  const int2 e = ij[c];
  atomicAdd(&cnt[e.x], 1);
  atomicAdd(&cnt[e.y], 1);
}

__global__ void k_inc_fill2(int nrec, int rec0, const int2* ij, const int* off, int* fill, int* ent)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrec) return;
  const int2 e = ij[c];
  const int rec = rec0 + c;
  ent[off[e.x] + atomicAdd(&fill[e.x], 1)] = rec << 1;
  ent[off[e.y] + atomicAdd(&fill[e.y], 1)] = (rec << 1) | 1;
}

// each particle's run in ascending entry order: pairs by key, then bonds
__global__ void k_inc_sort2(int n, const int* off, int* ent)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = off[i], b = off[i + 1];
  for (int s = a + 1; s < b; s++) {
    const int v = ent[s];
    int j = s - 1;
    while (j >= a && ent[j] > v) {
      ent[j + 1] = ent[j];
      j--;
    }
    ent[j + 1] = v;
  }
}


// ---------------------------------------------------------------------------
// adjoint kernels
// ---------------------------------------------------------------------------
// lam: per particle 4 double4 (x, v, w, q) holding the adjoint of the state
// after the step on entry and the partial adjoint of the state before it on
// exit; lfm: per particle (lambda_F, lambda_M) for the interaction adjoints.
// coef[s]: d loss_s / d stress_s / area_s at a sample step (else 0).
__global__ void k_adj_particle(DState Sn, DState Sn1, double4* lam, double* lfm, DStep sp,
                               const double* coef)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Sn.n) return;
  double4* L = lam + 4 * (size_t)i;
  const double4 Lx = L[0], Lv1 = L[1], Lw1 = L[2], Lq1 = L[3];
  const double dt = sp.dt, h = dt / 2.0;
  const double4 Vv = Sn.v[i], A = Sn.w[i], Q = Sn.q[i];
  const double4 A1 = Sn1.w[i], Q1 = Sn1.q[i];
  // x' = x + v' dt, v' = v + F invm dt
  const dv3 lv1 = add3(mk3(Lv1.x, Lv1.y, Lv1.z), mul3(mk3(Lx.x, Lx.y, Lx.z), dt));
  const dv3 lF = mul3(lv1, Vv.w * dt);
  dv3 lx = mk3(Lx.x, Lx.y, Lx.z);
  // q' = qt / |qt|, qt = q + h (0, w') q, w' = w + M invI dt
  const dv3 w1 = mk3(A1.x, A1.y, A1.z);
  const dquat q = dquat{Q.x, mk3(Q.y, Q.z, Q.w)};
  const dquat dq = qmul4(dquat{0, w1}, q);
  const dquat qt = dquat{q.s + h * dq.s, add3(q.v, mul3(dq.v, h))};
  const double mod = sqrt(qt.s * qt.s + qt.v.x * qt.v.x + qt.v.y * qt.v.y + qt.v.z * qt.v.z);
  const dquat qn = dquat{Q1.x, mk3(Q1.y, Q1.z, Q1.w)};
  const double pr = qn.s * Lq1.x + qn.v.x * Lq1.y + qn.v.y * Lq1.z + qn.v.z * Lq1.w;
  const dquat lqt = dquat{(Lq1.x - qn.s * pr) / mod,
                          div3(sub3(mk3(Lq1.y, Lq1.z, Lq1.w), mul3(qn.v, pr)), mod)};
  // d(dq)/dq^T lqt = (w1 . lqt.v, -w1 lqt.s - w1 x lqt.v)
  const dquat lq = dquat{lqt.s + h * dot3(w1, lqt.v),
                         add3(lqt.v, mul3(sub3(mul3(w1, -lqt.s), cross3(w1, lqt.v)), h))};
  // d(dq)/dw1^T lqt = -q.v lqt.s + q.s lqt.v + q.v x lqt.v
  const dv3 lw1 = add3(mk3(Lw1.x, Lw1.y, Lw1.z),
                       mul3(add3(add3(mul3(q.v, -lqt.s), mul3(lqt.v, q.s)), cross3(q.v, lqt.v)), h));
  const dv3 lM = mul3(lw1, dt * A.w);
  // walls: f = n k' (r - (x - o).n) while touching; the measured wall's normal
  // force also enters the stress
  const double4 P = Sn.x[i];
  const double r = P.w;
  const double cf = coef ? coef[Sn.spec[i]] : 0.0;
  for (int k = 0; k < sp.nw; k++) {
    const DWall& W = sp.wall[k];
    const dv3 n = mk3(W.n[0], W.n[1], W.n[2]);
    const double dist = dot3(sub3(mk3(P.x, P.y, P.z), mk3(W.o[0], W.o[1], W.o[2])), n);
    if (dist < r) {
      const double kk = (3.141592654 * r) * W.k;
      // d f / d x = -kk n n^T;  d (f . n) / d x = -kk n
      double s = dot3(n, lF);
      if (W.measure) s += cf;
      lx = sub3(lx, mul3(n, kk * s));
    }
  }
  L[0] = make_double4(lx.x, lx.y, lx.z, 0.0);
  L[1] = make_double4(lv1.x, lv1.y, lv1.z, 0.0);
  L[2] = make_double4(lw1.x, lw1.y, lw1.z, 0.0);
  L[3] = make_double4(lq.s, lq.v.x, lq.v.y, lq.v.z);
  double* o = lfm + 6 * (size_t)i;
  o[0] = lF.x; o[1] = lF.y; o[2] = lF.z;
  o[3] = lM.x; o[4] = lM.y; o[5] = lM.z;
}

#define NB_TAN 16
#define NF_TAN 24

// bond law VJP: c[k] = lambda_out . d out / d in_k over the inputs
// (x1, x2, q1, q2, theta_kn, theta_ks), one thread per bond, by the law's
// hand-written reverse mode (tbond_law_vjp) on the branch the forward step took
#define ADJ_BOND_BS 256
__global__ __launch_bounds__(ADJ_BOND_BS) void k_adj_bond(DState S, DBonds B,
                                                           const double4* theta, int step,
                                                           const double* lfm, double* ab)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.nb) return;
  double* c = ab + NB_TAN * (size_t)b;
  if (B.bstep[b] < step) return;  // broken earlier: no record (k_adj_gather skips it)
  const int2 e = B.ij[b];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y], Q1 = S.q[e.x], Q2 = S.q[e.y];
  const double4 D = B.D0[b];
  const int s = S.spec[e.x];
  const double4 th = theta[s];
  const double2 tk = B.tkb[s];
  const double* l1 = lfm + 6 * (size_t)e.x;
  const double* l2 = lfm + 6 * (size_t)e.y;
  double kr, ks, kt, kb, dkr, dks, dkt, dkb;
  tbond_k<double>(th.x, th.y, tk.x, tk.y, D.w, P1.w, P2.w, 1, 1, 0, kr, ks, kt, kb);
  // d kr / d theta_kn, d ks / d theta_ks (the constants are linear in them)
  tbond_k<double>(1.0, 1.0, tk.x, tk.y, D.w, P1.w, P2.w, 1, 1, 0, dkr, dks, dkt, dkb);
  tv3<double> g1, g2;
  tquat<double> gq1, gq2;
  double gkr, gks;
  // f2 = -f1: particle 2's force adjoint enters with the opposite sign
  tbond_law_vjp(dx3(P1), tquat<double>{Q1.x, {Q1.y, Q1.z, Q1.w}}, P1.w, dx3(P2),
                tquat<double>{Q2.x, {Q2.y, Q2.z, Q2.w}}, P2.w, kr, ks, kt, kb,
                tv3<double>{D.x, D.y, D.z}, tv3<double>{l1[0] - l2[0], l1[1] - l2[1], l1[2] - l2[2]},
                tv3<double>{l1[3], l1[4], l1[5]}, tv3<double>{l2[3], l2[4], l2[5]}, g1, g2, gq1, gq2,
                gkr, gks);
  c[0] = g1.x; c[1] = g1.y; c[2] = g1.z;
  c[3] = g2.x; c[4] = g2.y; c[5] = g2.z;
  c[6] = gq1.s; c[7] = gq1.v.x; c[8] = gq1.v.y; c[9] = gq1.v.z;
  c[10] = gq2.s; c[11] = gq2.v.x; c[12] = gq2.v.y; c[13] = gq2.v.z;
  c[14] = gkr * dkr;
  c[15] = gks * dks;
}

// friction law VJP over the inputs (x1, x2, v1, v2, w1, w2, history,
// theta_kn, theta_ks, theta_mu), one thread per pair, by the law's hand-written
// reverse mode (tfriction_law_vjp) on the slip branch the forward step took;
// lh: adjoint of the history this step wrote; lh_prev: adjoint of the history
// it read
// act[c]: the record of pair c holds a (nonzero) Jacobian this step; the
// gather reads only such records.
__global__ void k_adj_fric(DState S, DCands C, const int* bstep, const double4* theta,
                           const double4* hin, int step, double dt, const double* lfm,
                           const double4* lh, double* ac, unsigned char* act, double4* lh_prev)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C.nc) return;
  double* o = ac + NF_TAN * (size_t)c;
  const int bb = C.bond[c];
  if (bb >= 0 && bstep[bb] >= step) {
    // excluded while bonded: zero Jacobian.  The history adjoint it would pass
    // back is not written: the pair was excluded in the step before too (a
    // bond that lives now lived then), so no step reads it.
    act[c] = 0;
    return;
  }
  const int2 e = C.ij[c];
  const double4 P1 = S.x[e.x], P2 = S.x[e.y];
  // apart (no force, history cleared): zero Jacobian
  const double dxx = P1.x - P2.x, dyy = P1.y - P2.y, dzz = P1.z - P2.z;
  const double eq = P1.w + P2.w;
  if (!((dxx * dxx + dyy * dyy + dzz * dzz) < eq * eq)) {
    act[c] = 0;
    lh_prev[c] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  act[c] = 1;
  const double4 V1 = S.v[e.x], V2 = S.v[e.y], W1 = S.w[e.x], W2 = S.w[e.y];
  const double4 th = theta[S.spec[e.x]];
  const double4 hh = hin[c];
  const double4 lo = lh[c];
  const double* l1 = lfm + 6 * (size_t)e.x;
  const double* l2 = lfm + 6 * (size_t)e.y;
  double kk, kks, dk, dks;
  tfriction_k<double>(th.x, th.y, 1, 1, P1.w, P2.w, 0, kk, kks);
  // d k / d theta_kn, d ks / d theta_ks (linear in them)
  tfriction_k<double>(1.0, 1.0, 1, 1, P1.w, P2.w, 0, dk, dks);
  tv3<double> g[8];
  // f2 = -f1: particle 2's force adjoint enters with the opposite sign
  tfriction_law_vjp(dx3(P1), dx3(V1), dx3(W1), P1.w, dx3(P2), dx3(V2), dx3(W2), P2.w, kk, kks, th.w,
                    dt, tv3<double>{hh.x, hh.y, hh.z}, (((int)hh.w) & 1) != 0,
                    tv3<double>{l1[0] - l2[0], l1[1] - l2[1], l1[2] - l2[2]},
                    tv3<double>{l1[3], l1[4], l1[5]}, tv3<double>{l2[3], l2[4], l2[5]},
                    tv3<double>{lo.x, lo.y, lo.z}, g);
  for (int t = 0; t < 7; t++) {
    o[3 * t] = g[t].x;
    o[3 * t + 1] = g[t].y;
    o[3 * t + 2] = g[t].z;
  }
  o[21] = g[7].x * dk;
  o[22] = g[7].y * dks;
  o[23] = g[7].z;
  lh_prev[c] = make_double4(g[6].x, g[6].y, g[6].z, 0.0);
}

// per particle: add the interactions' adjoints (CSR order) to the state
// adjoint; parameter gradients of the interactions whose first particle this
// is go to gth (kn, ks, -, mu).  Entries go ADJ_GATHER_CH at a time with their
// loads in flight together; records without a Jacobian this step (pairs apart
// or excluded, bonds broken earlier) are skipped: they would add zeros.
__global__ void k_adj_gather(int n, const int* inc_off, const int* inc_ent, int ncand,
                             const double* ac, const double* ab, const unsigned char* cact,
                             const int* bstep, int step, double4* lam, double4* gth)
{
  constexpr int ADJ_GATHER_RUN = 12, ADJ_GATHER_CH = 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4* L = lam + 4 * (size_t)i;
  double4 lx = L[0], lv = L[1], lw = L[2], lq = L[3];
  double4 g = gth[i];
  const int e0 = inc_off[i], e1 = inc_off[i + 1];
  for (int r0 = e0; r0 < e1; r0 += ADJ_GATHER_RUN) {
    int ent[ADJ_GATHER_RUN];
#pragma unroll
    for (int k = 0; k < ADJ_GATHER_RUN; k++) ent[k] = r0 + k < e1 ? inc_ent[r0 + k] : -1;
    unsigned act = 0;
#pragma unroll
    for (int k = 0; k < ADJ_GATHER_RUN; k++) {
      const int rec = ent[k] >> 1;
      const bool on =
          ent[k] >= 0 && (rec < ncand ? cact[rec] != 0 : bstep[rec - ncand] >= step);
      act |= (on ? 1u : 0u) << k;
    }
#pragma unroll
    for (int c0 = 0; c0 < ADJ_GATHER_RUN; c0 += ADJ_GATHER_CH) {
      if (!((act >> c0) & ((1u << ADJ_GATHER_CH) - 1u))) continue;
      // per entry: x (3), v or q (4), w (3), parameter terms (3)
      double a[ADJ_GATHER_CH][13];
#pragma unroll
      for (int k = 0; k < ADJ_GATHER_CH; k++) {
        if (!((act >> (c0 + k)) & 1u)) continue;
        const int rec = ent[c0 + k] >> 1, side = ent[c0 + k] & 1;
        if (rec < ncand) {
          const double* c = ac + NF_TAN * (size_t)rec;
          const int xo = side ? 3 : 0, vo = side ? 9 : 6, wo = side ? 15 : 12;
          a[k][0] = c[xo]; a[k][1] = c[xo + 1]; a[k][2] = c[xo + 2];
          a[k][3] = c[vo]; a[k][4] = c[vo + 1]; a[k][5] = c[vo + 2]; a[k][6] = 0.0;
          a[k][7] = c[wo]; a[k][8] = c[wo + 1]; a[k][9] = c[wo + 2];
          if (!side) {
            a[k][10] = c[21]; a[k][11] = c[22]; a[k][12] = c[23];
          }
        } else {
          const double* c = ab + NB_TAN * (size_t)(rec - ncand);
          const int xo = side ? 3 : 0, qo = side ? 10 : 6;
          a[k][0] = c[xo]; a[k][1] = c[xo + 1]; a[k][2] = c[xo + 2];
          a[k][3] = c[qo]; a[k][4] = c[qo + 1]; a[k][5] = c[qo + 2]; a[k][6] = c[qo + 3];
          if (!side) {
            a[k][10] = c[14]; a[k][11] = c[15];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < ADJ_GATHER_CH; k++) {
        if (!((act >> (c0 + k)) & 1u)) continue;
        const int rec = ent[c0 + k] >> 1, side = ent[c0 + k] & 1;
        lx.x += a[k][0]; lx.y += a[k][1]; lx.z += a[k][2];
        if (rec < ncand) {
          lv.x += a[k][3]; lv.y += a[k][4]; lv.z += a[k][5];
          lw.x += a[k][7]; lw.y += a[k][8]; lw.z += a[k][9];
          if (!side) {
            g.x += a[k][10];
            g.y += a[k][11];
            g.w += a[k][12];
          }
        } else {
          lq.x += a[k][3]; lq.y += a[k][4]; lq.z += a[k][5]; lq.w += a[k][6];
          if (!side) {
            g.x += a[k][10];
            g.y += a[k][11];
          }
        }
      }
    }
  }
  L[0] = lx; L[1] = lv; L[2] = lw; L[3] = lq;
  gth[i] = g;
}

// adjoint of the history read at the first step of an epoch, to the pairs of
// the previous epoch it was taken from
__global__ void k_map_lh(int nc, const int* src, const double4* lnew, double4* lold)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int s = src[c];
  if (s >= 0) lold[s] = lnew[c];
}

__global__ void k_spec_sum4(const double4* val, const int* spec_start, double4* out)
{
  __shared__ double4 red[1024];
  const int s = blockIdx.x;
  const int a = spec_start[s], b = spec_start[s + 1];
  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  for (int i = a + (int)threadIdx.x; i < b; i += blockDim.x) {
    const double4 v = val[i];
    acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z; acc.w = acc.w + v.w;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double4 t = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = 0; k < (int)blockDim.x; k++) {
      t.x = t.x + red[k].x; t.y = t.y + red[k].y; t.z = t.z + red[k].z; t.w = t.w + red[k].w;
    }
    out[s] = t;
  }
}

inline int nblocks(long long n, int bs) { return (int)std::max<long long>(1, (n + bs - 1) / bs); }

// one neighbour-table epoch: the pairs found at a rebuild and the CSR of every
// particle's interactions (pairs, then bonds)
struct Epoch {
  int start = 0;  // first step it serves
  int nc = 0;
  DevBuf<int2> ij;
  DevBuf<long long> key;
  DevBuf<int> bond, src;
  DevBuf<int> inc_off, inc_ent;
};

struct Checkpoint {
  int step = 0;
  int epoch = 0;
  DevBuf<double4> x, v, w, q, o, hist;
  DevBuf<int> bstep;
};

}  // namespace

struct esp_diff {
  esp_diff_config cfg;
  GridParams grid;
  int ncell = 0;
  hipStream_t stream = nullptr;
  int bs = 256;
  std::string err;
  // particles
  int n = 0, nspec = 0;
  std::vector<int> h_spec, h_spec_start;
  std::vector<double> h_pos, h_rad, h_mass, h_vel;
  DevBuf<double4> x, v, w, q, o;
  DevBuf<int> spec, spec_start, id;
  DevBuf<unsigned> flags;
  // bonds (sorted by key)
  int nb = 0;
  std::vector<int> h_bi, h_bj;
  std::vector<double> h_rel, h_tk, h_bk;
  DevBuf<int2> bij;
  DevBuf<long long> bkey;
  DevBuf<double4> bD0, brel;
  DevBuf<int> bstep;
  DevBuf<double2> tkb;
  // walls, gravity, run
  int nw = 0;
  DWall wall0[DIFF_MAX_WALLS];
  double wstep[DIFF_MAX_WALLS][3];
  std::vector<double> h_area;
  DevBuf<double> area;
  int has_g = 0;
  double g[3] = {0, 0, 0};
  int nsteps = 0, every = 1;
  // run state
  std::vector<std::unique_ptr<Epoch>> epochs;
  std::vector<std::unique_ptr<Checkpoint>> ckpts;
  std::vector<int> rebuild_steps;
  DevBuf<double4> H[2];  // shear histories: step s reads H[s & 1], writes the other
  DevBuf<DRun> run;
  DevBuf<double> wo;     // wall origins per step of the run [step][wall][3]
  // the forward steps of one epoch captured as a graph (device-driven steps)
  hipGraphExec_t gexec = nullptr;
  int g_epoch = -1, g_len = 0;
  bool use_graph = true;
  int graph_steps = 16;
  DevBuf<double> bout, cout, wfn, stress_dev;
  DevBuf<unsigned char> cact;  // per pair of the epoch: its record holds this step's force
  DevBuf<double4> theta;
  std::vector<double> h_theta, h_target, h_stress;
  DevBuf<unsigned long long> md;
  DevBuf<unsigned long long> digest;  // per step: branch digest of the last forward run
  // branch digests are computed only when asked for (esp_diff_track_branches):
  // every wave of the contact and gather kernels adds to one counter per step
  bool track_digest = false;
  // the state and break steps at the end of the last forward run: the backward
  // sweep recomputes earlier states in the same arrays and puts these back
  DevBuf<double4> fx, fv, fw, fq;
  DevBuf<int> fbstep;
  // tangent-linear run (forward with a parameter direction)
  bool tan_on = false;
  DevBuf<double4> tx, tv, tw, tq, TH[2], dtheta;
  DevBuf<double> bt, ct, wft, stress_t;
  std::vector<double> h_stress_t;
  DTan tan() { return DTan{tx.p, tv.p, tw.p, tq.p}; }
  DevBuf<int> cell, cell_count, cell_start, cell_fill, cell_items, ccnt, coff, cp, cq, scan_tmp,
      inc_cnt, inc_fill;
  DevBuf<long long> ckey;
  bool have_forward = false;
  double t_fwd = 0, t_bwd = 0;

  DState state()
  {
    return DState{x.p, v.p, w.p, q.p, spec.p, n};
  }
  // wall origins of a step: moved by their step before every step (the
  // pre-time-step moveWallBy of the main engine's setWallStepMotion)
  DStep step_params(int step)
  {
    DStep s;
    memset(&s, 0, sizeof s);
    s.nw = nw;
    for (int k = 0; k < nw; k++) {
      s.wall[k] = wall0[k];
      for (int t = 0; t < step; t++)
        for (int c = 0; c < 3; c++) s.wall[k].o[c] = s.wall[k].o[c] + wstep[k][c];
    }
    s.has_g = has_g;
    for (int c = 0; c < 3; c++) s.g[c] = g[c];
    s.dt = cfg.dt;
    s.step = step;
    return s;
  }
};

namespace {

void dsync(esp_diff* d) { DCHK(hipStreamSynchronize(d->stream)); }

int dscan(esp_diff* d, const int* in, int* out, int cnt)
{
  d->scan_tmp.reserve(scan_tmp_size(cnt));
  launch_exclusive_scan(d->stream, in, out, cnt, d->scan_tmp.p);
  DCHK(hipGetLastError());
  int total = 0;
  DCHK(hipMemcpyAsync(&total, out + cnt, sizeof(int), hipMemcpyDeviceToHost, d->stream));
  dsync(d);
  return total;
}

// neighbour search (NeighborTable::build + getFullList) -> a new epoch
void rebuild(esp_diff* d, int step)
{
  const int n = d->n;
  PView pv;
  memset(&pv, 0, sizeof pv);
  pv.pos = d->x.p; pv.vel = d->v.p; pv.ang = d->w.p; pv.quat = d->q.p; pv.old = d->o.p;
  pv.flags = d->flags.p; pv.id = d->id.p; pv.n = n;
  DCHK(hipMemsetAsync(d->cell_count.p, 0, sizeof(int) * (d->ncell + 1), d->stream));
  DCHK(hipMemsetAsync(d->md.p, 0, sizeof(unsigned long long) * 2, d->stream));
  int* nout = reinterpret_cast<int*>(d->md.p + 1);
  launch_cells_count(d->stream, pv, d->grid, d->cell.p, d->cell_count.p, nout);
  int hout = 0;
  DCHK(hipMemcpyAsync(&hout, nout, sizeof(int), hipMemcpyDeviceToHost, d->stream));
  dsync(d);
  if (hout > 0)
    throw DiffError(ESP_ERR_STATE, "particles left the neighbour-table domain (the differentiable "
                                   "engine keeps every particle)");
  dscan(d, d->cell_count.p, d->cell_start.p, d->ncell);
  DCHK(hipMemsetAsync(d->cell_fill.p, 0, sizeof(int) * d->ncell, d->stream));
  launch_cells_fill(d->stream, n, d->cell.p, d->cell_start.p, d->cell_fill.p, d->cell_items.p);
  launch_cells_sort(d->stream, d->ncell, d->cell_start.p, d->cell_items.p);
  launch_pairs(d->stream, 0, 0, pv, d->grid, d->cfg.verlet_dist, d->cell.p, d->cell_start.p,
               d->cell_items.p, d->ccnt.p, nullptr, nullptr, nullptr, nullptr);
  const int nc = dscan(d, d->ccnt.p, d->coff.p, n);
  d->cp.reserve(std::max(nc, 1));
  d->cq.reserve(std::max(nc, 1));
  d->ckey.reserve(std::max(nc, 1));
  launch_pairs(d->stream, 1, 0, pv, d->grid, d->cfg.verlet_dist, d->cell.p, d->cell_start.p,
               d->cell_items.p, nullptr, d->coff.p, d->cp.p, d->cq.p, d->ckey.p);
  DCHK(hipGetLastError());
  auto E = std::make_unique<Epoch>();
  E->start = step;
  E->nc = nc;
  const size_t c1 = std::max(nc, 1);
  E->ij.reserve(c1);
  E->key.reserve(c1);
  E->bond.reserve(c1);
  E->src.reserve(c1);
  DCHK(hipMemcpyAsync(E->key.p, d->ckey.p, sizeof(long long) * nc, hipMemcpyDeviceToDevice,
                      d->stream));
  const Epoch* P = d->epochs.empty() ? nullptr : d->epochs.back().get();
  const int bs = d->bs;
  if (nc > 0)
    hipLaunchKernelGGL(k_epoch_pairs, dim3(nblocks(nc, bs)), dim3(bs), 0, d->stream, nc, d->cp.p,
                       d->cq.p, d->ckey.p, d->bkey.p, d->nb, P ? P->key.p : nullptr, P ? P->nc : 0,
                       E->ij.p, E->bond.p, E->src.p);
  // CSR: pairs (records 0..nc-1), then bonds (nc..nc+nb-1)
  d->inc_cnt.reserve(n + 1);
  d->inc_fill.reserve(n + 1);
  DCHK(hipMemsetAsync(d->inc_cnt.p, 0, sizeof(int) * (n + 1), d->stream));
  if (nc > 0)
    hipLaunchKernelGGL(k_inc_count2, dim3(nblocks(nc, bs)), dim3(bs), 0, d->stream, nc, E->ij.p,
                       d->inc_cnt.p);
  if (d->nb > 0)
    hipLaunchKernelGGL(k_inc_count2, dim3(nblocks(d->nb, bs)), dim3(bs), 0, d->stream, d->nb,
                       d->bij.p, d->inc_cnt.p);
  E->inc_off.reserve(n + 1);
  const int tot = dscan(d, d->inc_cnt.p, E->inc_off.p, n);
  E->inc_ent.reserve(std::max(tot, 1));
  DCHK(hipMemsetAsync(d->inc_fill.p, 0, sizeof(int) * (n + 1), d->stream));
  if (nc > 0)
    hipLaunchKernelGGL(k_inc_fill2, dim3(nblocks(nc, bs)), dim3(bs), 0, d->stream, nc, 0, E->ij.p,
                       E->inc_off.p, d->inc_fill.p, E->inc_ent.p);
  if (d->nb > 0)
    hipLaunchKernelGGL(k_inc_fill2, dim3(nblocks(d->nb, bs)), dim3(bs), 0, d->stream, d->nb, nc,
                       d->bij.p, E->inc_off.p, d->inc_fill.p, E->inc_ent.p);
  hipLaunchKernelGGL(k_inc_sort2, dim3(nblocks(n, bs)), dim3(bs), 0, d->stream, n, E->inc_off.p,
                     E->inc_ent.p);
  // shear histories carried by key (PIS_ED::update: persistent pairs keep theirs;
  // a pair that is not is apart and holds zero anyway)
  // (the step reads buffer step & 1)
  const int cur = step & 1;
  DevBuf<double4> hn;
  hn.reserve(c1);
  if (nc > 0) {
    hipLaunchKernelGGL(k_carry_hist, dim3(nblocks(nc, bs)), dim3(bs), 0, d->stream, nc, E->src.p,
                       P ? d->H[cur].p : nullptr, hn.p);
    if (!P) DCHK(hipMemsetAsync(hn.p, 0, sizeof(double4) * nc, d->stream));
  }
  DevBuf<double4> tn;
  if (d->tan_on) {
    tn.reserve(c1);
    if (nc > 0) {
      hipLaunchKernelGGL(k_carry_hist, dim3(nblocks(nc, bs)), dim3(bs), 0, d->stream, nc, E->src.p,
                         P ? d->TH[cur].p : nullptr, tn.p);
      if (!P) DCHK(hipMemsetAsync(tn.p, 0, sizeof(double4) * nc, d->stream));
    }
    d->ct.reserve(c1 * 9);
  }
  DCHK(hipGetLastError());
  dsync(d);
  d->H[cur].swap(hn);
  d->H[cur ^ 1].reserve(c1);
  // both buffers equal: a pair flagged "zero in both buffers" writes neither
  if (nc > 0)
    DCHK(hipMemcpyAsync(d->H[cur ^ 1].p, d->H[cur].p, sizeof(double4) * nc,
                        hipMemcpyDeviceToDevice, d->stream));
  if (d->tan_on) {
    d->TH[cur].swap(tn);
    d->TH[cur ^ 1].reserve(c1);
  }
  d->epochs.push_back(std::move(E));
  d->rebuild_steps.push_back(step);
  d->bout.reserve((size_t)std::max(d->nb, 1) * 9);
  d->cout.reserve(c1 * 9);
  d->cact.reserve(c1);
  if (d->tan_on) d->bt.reserve((size_t)std::max(d->nb, 1) * 9);
}

template <class T>
void d2d(esp_diff* d, T* dst, const T* src, size_t cnt)
{
  if (cnt) DCHK(hipMemcpyAsync(dst, src, cnt * sizeof(T), hipMemcpyDeviceToDevice, d->stream));
}

void save_checkpoint(esp_diff* d, int step)
{
  auto C = std::make_unique<Checkpoint>();
  C->step = step;
  C->epoch = (int)d->epochs.size() - 1;
  const size_t n = std::max(d->n, 1);
  C->x.reserve(n); C->v.reserve(n); C->w.reserve(n); C->q.reserve(n); C->o.reserve(n);
  d2d(d, C->x.p, d->x.p, d->n); d2d(d, C->v.p, d->v.p, d->n); d2d(d, C->w.p, d->w.p, d->n);
  d2d(d, C->q.p, d->q.p, d->n); d2d(d, C->o.p, d->o.p, d->n);
  const int nc = d->epochs.back()->nc;
  C->hist.reserve(std::max(nc, 1));
  d2d(d, C->hist.p, d->H[step & 1].p, nc);
  C->bstep.reserve(std::max(d->nb, 1));
  d2d(d, C->bstep.p, d->bstep.p, d->nb);
  d->ckpts.push_back(std::move(C));
}

void restore_checkpoint(esp_diff* d, const Checkpoint& C)
{
  d2d(d, d->x.p, C.x.p, d->n); d2d(d, d->v.p, C.v.p, d->n); d2d(d, d->w.p, C.w.p, d->n);
  d2d(d, d->q.p, C.q.p, d->n); d2d(d, d->o.p, C.o.p, d->n);
  const int nc = d->epochs[C.epoch]->nc;
  d->H[0].reserve(std::max(nc, 1));
  d->H[1].reserve(std::max(nc, 1));
  // the history the checkpoint's step reads, in both buffers (a pair flagged
  // "zero in both buffers" writes neither)
  d2d(d, d->H[0].p, C.hist.p, nc);
  d2d(d, d->H[1].p, C.hist.p, nc);
  d2d(d, d->bstep.p, C.bstep.p, d->nb);
}

void upload_initial(esp_diff* d)
{
  const int n = d->n;
  std::vector<double4> X(n), V(n), W(n), Q(n), O(n);
  for (int i = 0; i < n; i++) {
    const double m = d->h_mass[i], r = d->h_rad[i];
    X[i] = make_double4(d->h_pos[3 * i], d->h_pos[3 * i + 1], d->h_pos[3 * i + 2], r);
    V[i] = make_double4(d->h_vel[3 * i], d->h_vel[3 * i + 1], d->h_vel[3 * i + 2],
                        m != 0.0 ? 1.0 / m : 0.0);
    const double I = 0.4 * m * r * r;  // CRotParticle, 3-D (RotParticle.cpp:80-124)
    W[i] = make_double4(0.0, 0.0, 0.0, 1.0 / I);
    Q[i] = make_double4(1.0, 0.0, 0.0, 0.0);
    O[i] = make_double4(X[i].x, X[i].y, X[i].z, m);
  }
  const size_t c = std::max(n, 1);
  d->x.reserve(c); d->v.reserve(c); d->w.reserve(c); d->q.reserve(c); d->o.reserve(c);
  DCHK(hipMemcpyAsync(d->x.p, X.data(), sizeof(double4) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->v.p, V.data(), sizeof(double4) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->w.p, W.data(), sizeof(double4) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->q.p, Q.data(), sizeof(double4) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->o.p, O.data(), sizeof(double4) * n, hipMemcpyHostToDevice, d->stream));
  std::vector<int> bs(std::max(d->nb, 1), INT_MAX);
  DCHK(hipMemcpyAsync(d->bstep.p, bs.data(), sizeof(int) * d->nb, hipMemcpyHostToDevice,
                      d->stream));
  dsync(d);
}

// one forward step (no search): forces, update, sample
// The launches of one forward step (no search): tangents (tangent runs),
// forces, update and sampling; the step index comes from d->run on the device
// (k_advance ends the step), so the same launches serve every step and can be
// captured in a graph.  record: branch digests into d->digest (forward runs);
// halting: the displacement test may stop the queue (not while recomputing a
// recorded segment).
void fwd_step(esp_diff* d, const Epoch& E, bool record, bool halting)
{
  const int bs = d->bs;
  DState S = d->state();
  DBonds B{d->nb, d->bij.p, d->bD0.p, d->brel.p, d->bstep.p, d->tkb.p};
  DCands C{E.nc, E.ij.p, E.bond.p};
  DRun* R = d->run.p;
  const DTan T = d->tan_on ? d->tan() : DTan{nullptr, nullptr, nullptr, nullptr};
  if (d->tan_on) {  // tangents from the state before the step (before the forward kernels)
    if (d->nb > 0)
      hipLaunchKernelGGL(k_bond_tan, dim3(nblocks(d->nb, bs)), dim3(bs), 0, d->stream, S, T, B,
                         d->theta.p, d->dtheta.p, R, d->bt.p);
    if (E.nc > 0)
      hipLaunchKernelGGL(k_fric_tan, dim3(nblocks(E.nc, bs)), dim3(bs), 0, d->stream, S, T, C,
                         d->bstep.p, d->theta.p, d->dtheta.p, d->H[0].p, d->H[1].p, d->TH[0].p,
                         d->TH[1].p, R, d->cfg.dt, d->ct.p);
  }
  if (d->nb > 0)
    hipLaunchKernelGGL(k_bond_fwd, dim3(nblocks(d->nb, bs)), dim3(bs), 0, d->stream, S, B,
                       d->theta.p, R, d->bout.p);
  if (E.nc > 0)
    hipLaunchKernelGGL(k_fric_fwd, dim3(nblocks(E.nc, bs)), dim3(bs), 0, d->stream, S, C, d->bstep.p,
                       d->theta.p, d->H[0].p, d->H[1].p, R, d->cfg.dt, d->cout.p, d->cact.p,
                       record && d->track_digest ? d->digest.p : nullptr);
  unsigned long long thresh = ~0ull;
  if (halting) {
    const double a = 0.5 * d->cfg.verlet_dist;
    const double sr = a * a;
    memcpy(&thresh, &sr, 8);
  }
  DStep sp = d->step_params(0);
  sp.wo = d->wo.p;
  hipLaunchKernelGGL(k_gather_fwd, dim3(nblocks(d->n, bs)), dim3(bs), 0, d->stream, S, d->o.p,
                     E.inc_off.p, E.inc_ent.p, E.nc, d->cout.p, d->bout.p, d->cact.p, d->bstep.p,
                     sp, d->every, d->wfn.p,
                     R, thresh, record && d->track_digest ? d->digest.p : nullptr, T, d->ct.p,
                     d->bt.p,
                     d->tan_on ? d->wft.p : nullptr);
  if (record) {
    hipLaunchKernelGGL(k_spec_sum, dim3(d->nspec), dim3(256), 0, d->stream, R, d->every, d->wfn.p,
                       d->spec_start.p, d->nspec, d->stress_dev.p);
    if (d->tan_on)
      hipLaunchKernelGGL(k_spec_sum, dim3(d->nspec), dim3(256), 0, d->stream, R, d->every,
                         d->wft.p, d->spec_start.p, d->nspec, d->stress_t.p);
  }
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, d->stream, R);
  DCHK(hipGetLastError());
}

void set_run(esp_diff* d, int step, int end)
{
  hipLaunchKernelGGL(k_set_run, dim3(1), dim3(64), 0, d->stream, d->run.p, step, 0, end);
  DCHK(hipGetLastError());
}

void drop_graph(esp_diff* d)
{
  if (d->gexec) {
    (void)hipStreamSynchronize(d->stream);
    (void)hipGraphExecDestroy(d->gexec);
  }
  d->gexec = nullptr;
  d->g_epoch = -1;
}

// forward steps d->run.step .. end on the current epoch (the run record is
// set): graph_steps forward steps captured once per epoch and replayed, or
// plain launches
void run_chunk(esp_diff* d, int from, int end)
{
  const Epoch& E = *d->epochs.back();
  const int G = std::max(1, d->graph_steps);
  if (!d->use_graph) {
    for (int s = from; s <= end; s++) fwd_step(d, E, true, true);
    return;
  }
  const int ep = (int)d->epochs.size() - 1;
  if (d->g_epoch != ep || d->g_len != G) {
    drop_graph(d);
    hipGraph_t g = nullptr;
    DCHK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < G; k++) fwd_step(d, E, true, true);
    DCHK(hipStreamEndCapture(d->stream, &g));
    const hipError_t rc = hipGraphInstantiate(&d->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    DCHK(rc);
    d->g_epoch = ep;
    d->g_len = G;
  }
  for (int s = from; s <= end; s += G) DCHK(hipGraphLaunch(d->gexec, d->stream));
}

int nsamples(const esp_diff* d) { return d->nsteps / d->every; }

// dtheta (nullable): a parameter direction per specimen; the run then also
// carries the state's tangent along it and dloss gets d loss / d epsilon at
// theta + epsilon dtheta
void forward(esp_diff* d, const double* theta, const double* target, double* stress, double* loss,
             const double* dtheta = nullptr, double* dloss = nullptr)
{
  if (d->n <= 0 || d->nsteps <= 0) throw DiffError(ESP_ERR_STATE, "particles and run not set");
  hipEvent_t e0, e1;
  DCHK(hipEventCreate(&e0));
  DCHK(hipEventCreate(&e1));
  DCHK(hipEventRecord(e0, d->stream));
  const int ns = nsamples(d);
  d->h_theta.assign(theta, theta + 4 * d->nspec);
  d->h_target.assign(target, target + (size_t)ns * d->nspec);
  d->theta.reserve(d->nspec);
  std::vector<double4> th(d->nspec);
  for (int s = 0; s < d->nspec; s++)
    th[s] = make_double4(theta[4 * s], theta[4 * s + 1], theta[4 * s + 2], theta[4 * s + 3]);
  DCHK(hipMemcpyAsync(d->theta.p, th.data(), sizeof(double4) * d->nspec, hipMemcpyHostToDevice,
                      d->stream));
  d->stress_dev.reserve((size_t)std::max(ns, 1) * d->nspec);
  d->digest.reserve((size_t)d->nsteps + 1);
  DCHK(hipMemsetAsync(d->digest.p, 0, sizeof(unsigned long long) * (d->nsteps + 1), d->stream));
  d->epochs.clear();
  d->ckpts.clear();
  d->rebuild_steps.clear();
  d->tan_on = dtheta != nullptr;
  if (d->tan_on) {
    const size_t c = std::max(d->n, 1);
    d->tx.reserve(c); d->tv.reserve(c); d->tw.reserve(c); d->tq.reserve(c);
    for (DevBuf<double4>* b : {&d->tx, &d->tv, &d->tw, &d->tq})
      DCHK(hipMemsetAsync(b->p, 0, sizeof(double4) * c, d->stream));  // the start state is fixed
    d->wft.reserve(c);
    d->bt.reserve((size_t)std::max(d->nb, 1) * 9);
    d->stress_t.reserve((size_t)std::max(ns, 1) * d->nspec);
    d->dtheta.reserve(d->nspec);
    std::vector<double4> dt4(d->nspec);
    for (int s = 0; s < d->nspec; s++)
      dt4[s] = make_double4(dtheta[4 * s], dtheta[4 * s + 1], dtheta[4 * s + 2], dtheta[4 * s + 3]);
    DCHK(hipMemcpyAsync(d->dtheta.p, dt4.data(), sizeof(double4) * d->nspec, hipMemcpyHostToDevice,
                        d->stream));
    dsync(d);  // the copy reads dt4, which ends with this block
  }
  upload_initial(d);
  // wall origins of every step: moved by their step before every step (the
  // pre-time-step moveWallBy of the main engine's setWallStepMotion)
  {
    const int nw = std::max(d->nw, 1);
    std::vector<double> wo((size_t)(d->nsteps + 2) * nw * 3, 0.0);
    double o[DIFF_MAX_WALLS][3];
    for (int k = 0; k < d->nw; k++)
      for (int c = 0; c < 3; c++) o[k][c] = d->wall0[k].o[c];
    for (int t = 0; t <= d->nsteps + 1; t++) {
      for (int k = 0; k < d->nw; k++)
        for (int c = 0; c < 3; c++) wo[((size_t)t * nw + k) * 3 + c] = o[k][c];
      for (int k = 0; k < d->nw; k++)
        for (int c = 0; c < 3; c++) o[k][c] = o[k][c] + d->wstep[k][c];
    }
    d->wo.reserve(wo.size());
    DCHK(hipMemcpyAsync(d->wo.p, wo.data(), sizeof(double) * wo.size(), hipMemcpyHostToDevice,
                        d->stream));
    dsync(d);  // the copy reads wo, which ends with this block
  }
  d->run.reserve(1);
  d->wfn.reserve(std::max(d->n, 1));
  drop_graph(d);
  // chunks of steps between checkpoints queued without host decisions; a step
  // whose displacement asks for a neighbour search halts the queue, the host
  // rebuilds the table before the next step and continues
  const int K = std::max(1, d->cfg.checkpoint_every);
  bool need = true;
  int step = 1;
  while (step <= d->nsteps) {
    if (need) {
      drop_graph(d);
      rebuild(d, step);
      need = false;
    }
    if ((step - 1) % K == 0) save_checkpoint(d, step);
    const int end = std::min(d->nsteps, ((step - 1) / K + 1) * K);
    set_run(d, step, end);
    run_chunk(d, step, end);
    DRun r;
    DCHK(hipMemcpyAsync(&r, d->run.p, sizeof r, hipMemcpyDeviceToHost, d->stream));
    dsync(d);
    if (r.halt != 0) {
      need = true;
      step = r.halt + 1;
    } else {
      step = end + 1;
    }
  }
  drop_graph(d);
  {
    const size_t c = std::max(d->n, 1);
    d->fx.reserve(c); d->fv.reserve(c); d->fw.reserve(c); d->fq.reserve(c);
    d->fbstep.reserve(std::max(d->nb, 1));
    d2d(d, d->fx.p, d->x.p, d->n); d2d(d, d->fv.p, d->v.p, d->n);
    d2d(d, d->fw.p, d->w.p, d->n); d2d(d, d->fq.p, d->q.p, d->n);
    d2d(d, d->fbstep.p, d->bstep.p, d->nb);
  }
  d->h_stress.assign((size_t)ns * d->nspec, 0.0);
  DCHK(hipMemcpyAsync(d->h_stress.data(), d->stress_dev.p, sizeof(double) * ns * d->nspec,
                      hipMemcpyDeviceToHost, d->stream));
  if (d->tan_on) {
    d->h_stress_t.assign((size_t)ns * d->nspec, 0.0);
    DCHK(hipMemcpyAsync(d->h_stress_t.data(), d->stress_t.p, sizeof(double) * ns * d->nspec,
                        hipMemcpyDeviceToHost, d->stream));
  }
  DCHK(hipEventRecord(e1, d->stream));
  dsync(d);
  float ms = 0;
  DCHK(hipEventElapsedTime(&ms, e0, e1));
  d->t_fwd = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int s = 0; s < d->nspec; s++) {
    double acc = 0.0;
    for (int k = 0; k < ns; k++) {
      const double st = d->h_stress[(size_t)k * d->nspec + s] / d->h_area[s];
      if (stress) stress[(size_t)s * ns + k] = st;
      const double df = st - d->h_target[(size_t)s * ns + k];
      acc = acc + df * df;
    }
    loss[s] = ns > 0 ? acc / ns : 0.0;
    if (d->tan_on && dloss) {
      double dacc = 0.0;
      for (int k = 0; k < ns; k++) {
        const double st = d->h_stress[(size_t)k * d->nspec + s] / d->h_area[s];
        const double dst = d->h_stress_t[(size_t)k * d->nspec + s] / d->h_area[s];
        dacc = dacc + 2.0 * (st - d->h_target[(size_t)s * ns + k]) * dst;
      }
      dloss[s] = ns > 0 ? dacc / ns : 0.0;
    }
  }
  d->tan_on = false;
  d->have_forward = true;
}

void backward(esp_diff* d, double* grad)
{
  if (!d->have_forward) throw DiffError(ESP_ERR_STATE, "backward needs a forward run first");
  hipEvent_t e0, e1;
  DCHK(hipEventCreate(&e0));
  DCHK(hipEventCreate(&e1));
  DCHK(hipEventRecord(e0, d->stream));
  const int n = d->n, bs = d->bs, ns = nsamples(d);
  const int K = std::max(1, d->cfg.checkpoint_every);
  const size_t c = std::max(n, 1);
  // d loss / d stress_s at each sample, divided by the area (the stress is the
  // specimen's wall force over its area)
  std::vector<double> coef((size_t)std::max(ns, 1) * d->nspec, 0.0);
  for (int s = 0; s < d->nspec; s++)
    for (int k = 0; k < ns; k++) {
      const double st = d->h_stress[(size_t)k * d->nspec + s] / d->h_area[s];
      coef[(size_t)k * d->nspec + s] = 2.0 * (st - d->h_target[(size_t)s * ns + k]) / ns / d->h_area[s];
    }
  DevBuf<double> coef_d;
  coef_d.reserve(coef.size());
  DCHK(hipMemcpyAsync(coef_d.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice,
                      d->stream));
  DevBuf<double4> lam, gth, lh, lh2;
  DevBuf<double> lfm, ab, ac;
  DevBuf<unsigned char> aact;  // per pair: its record holds a Jacobian this step
  lam.reserve(4 * c);
  gth.reserve(c);
  lfm.reserve(6 * c);
  ab.reserve((size_t)std::max(d->nb, 1) * NB_TAN);
  DCHK(hipMemsetAsync(lam.p, 0, sizeof(double4) * 4 * c, d->stream));
  DCHK(hipMemsetAsync(gth.p, 0, sizeof(double4) * c, d->stream));
  // tape of one segment: states of steps s0..s1, histories read at s0..s1-1
  DevBuf<double4> tape;
  tape.reserve((size_t)(K + 1) * 4 * c);
  std::vector<std::unique_ptr<DevBuf<double4>>> htape;
  int lh_epoch = -1;  // epoch whose pairs lh is indexed by
  DBonds B{d->nb, d->bij.p, d->bD0.p, d->brel.p, d->bstep.p, d->tkb.p};
  auto epoch_of = [&](int step) {
    int e = 0;
    while (e + 1 < (int)d->epochs.size() && d->epochs[e + 1]->start <= step) e++;
    return e;
  };
  for (int seg = (int)d->ckpts.size() - 1; seg >= 0; seg--) {
    const Checkpoint& C = *d->ckpts[seg];
    const int s0 = C.step, s1 = std::min(s0 + K, d->nsteps + 1);
    restore_checkpoint(d, C);
    // recompute the segment forward, recording the states and histories
    if ((int)htape.size() < K) htape.resize(K);
    int cur = C.epoch;
    set_run(d, s0, s1 - 1);
    for (int step = s0; step < s1; step++) {
      const int e = epoch_of(step);
      if (e != cur) {  // the rebuild of this step: histories carried by key
        const Epoch& E = *d->epochs[e];
        DevBuf<double4> hn;
        hn.reserve(std::max(E.nc, 1));
        if (E.nc > 0)
          hipLaunchKernelGGL(k_carry_hist, dim3(nblocks(E.nc, bs)), dim3(bs), 0, d->stream, E.nc,
                             E.src.p, d->H[step & 1].p, hn.p);
        d->H[step & 1].swap(hn);
        d->H[(step + 1) & 1].reserve(std::max(E.nc, 1));
        d2d(d, d->H[(step + 1) & 1].p, d->H[step & 1].p, E.nc);
        d->cout.reserve((size_t)std::max(E.nc, 1) * 9);
        d->cact.reserve((size_t)std::max(E.nc, 1));
        cur = e;
      }
      const int t = step - s0;
      double4* T = tape.p + (size_t)t * 4 * c;
      d2d(d, T, d->x.p, n); d2d(d, T + c, d->v.p, n); d2d(d, T + 2 * c, d->w.p, n);
      d2d(d, T + 3 * c, d->q.p, n);
      const int nc = d->epochs[cur]->nc;
      if (!htape[t]) htape[t] = std::make_unique<DevBuf<double4>>();
      htape[t]->reserve(std::max(nc, 1));
      d2d(d, htape[t]->p, d->H[step & 1].p, nc);
      fwd_step(d, *d->epochs[cur], false, false);
    }
    {
      double4* T = tape.p + (size_t)(s1 - s0) * 4 * c;
      d2d(d, T, d->x.p, n); d2d(d, T + c, d->v.p, n); d2d(d, T + 2 * c, d->w.p, n);
      d2d(d, T + 3 * c, d->q.p, n);
    }
    // reverse sweep
    for (int step = s1 - 1; step >= s0; step--) {
      const int e = epoch_of(step);
      const Epoch& E = *d->epochs[e];
      if (lh_epoch != e) {
        // the history adjoint of the next step's pairs, taken back to this
        // step's pairs (the next step started an epoch) or started at zero
        DevBuf<double4> nl;
        nl.reserve(std::max(E.nc, 1));
        DCHK(hipMemsetAsync(nl.p, 0, sizeof(double4) * std::max(E.nc, 1), d->stream));
        if (lh_epoch >= 0) {
          const Epoch& N = *d->epochs[lh_epoch];
          if (N.nc > 0)
            hipLaunchKernelGGL(k_map_lh, dim3(nblocks(N.nc, bs)), dim3(bs), 0, d->stream, N.nc,
                               N.src.p, lh.p, nl.p);
        }
        lh.swap(nl);
        lh_epoch = e;
        lh2.reserve(std::max(E.nc, 1));
        ac.reserve((size_t)std::max(E.nc, 1) * NF_TAN);
        aact.reserve((size_t)std::max(E.nc, 1));
      }
      const int t = step - s0;
      double4* T0 = tape.p + (size_t)t * 4 * c;
      double4* T1 = tape.p + (size_t)(t + 1) * 4 * c;
      DState Sn{T0, T0 + c, T0 + 2 * c, T0 + 3 * c, d->spec.p, n};
      DState Sn1{T1, T1 + c, T1 + 2 * c, T1 + 3 * c, d->spec.p, n};
      const bool sample = step % d->every == 0;
      const double* cf = sample ? coef_d.p + (size_t)(step / d->every - 1) * d->nspec : nullptr;
      hipLaunchKernelGGL(k_adj_particle, dim3(nblocks(n, bs)), dim3(bs), 0, d->stream, Sn, Sn1,
                         lam.p, lfm.p, d->step_params(step), cf);
      if (d->nb > 0)
        hipLaunchKernelGGL(k_adj_bond, dim3(nblocks(d->nb, ADJ_BOND_BS)), dim3(ADJ_BOND_BS), 0,
                           d->stream, Sn, B, d->theta.p, step, lfm.p, ab.p);
      DCands Cd{E.nc, E.ij.p, E.bond.p};
      if (E.nc > 0)
        hipLaunchKernelGGL(k_adj_fric, dim3(nblocks(E.nc, bs)), dim3(bs), 0, d->stream, Sn, Cd,
                           d->bstep.p, d->theta.p, htape[t]->p, step, d->cfg.dt, lfm.p, lh.p, ac.p,
                           aact.p, lh2.p);
      hipLaunchKernelGGL(k_adj_gather, dim3(nblocks(n, bs)), dim3(bs), 0, d->stream, n,
                         E.inc_off.p, E.inc_ent.p, E.nc, ac.p, ab.p, aact.p, d->bstep.p, step,
                         lam.p, gth.p);
      DCHK(hipGetLastError());
      lh.swap(lh2);
    }
    dsync(d);
  }
  // the forward run's final state and break steps, for later queries
  d2d(d, d->x.p, d->fx.p, n); d2d(d, d->v.p, d->fv.p, n);
  d2d(d, d->w.p, d->fw.p, n); d2d(d, d->q.p, d->fq.p, n);
  d2d(d, d->bstep.p, d->fbstep.p, d->nb);
  DevBuf<double4> gs;
  gs.reserve(std::max(d->nspec, 1));
  hipLaunchKernelGGL(k_spec_sum4, dim3(d->nspec), dim3(256), 0, d->stream, gth.p, d->spec_start.p,
                     gs.p);
  std::vector<double4> hg(d->nspec);
  DCHK(hipMemcpyAsync(hg.data(), gs.p, sizeof(double4) * d->nspec, hipMemcpyDeviceToHost,
                      d->stream));
  DCHK(hipEventRecord(e1, d->stream));
  dsync(d);
  float ms = 0;
  DCHK(hipEventElapsedTime(&ms, e0, e1));
  d->t_bwd = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int s = 0; s < d->nspec; s++) {
    grad[4 * s] = hg[s].x;
    grad[4 * s + 1] = hg[s].y;
    grad[4 * s + 2] = 0.0;  // cohesion moves only the break events
    grad[4 * s + 3] = hg[s].w;
  }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
#define DAPI_BEGIN try {
#define DAPI_END                     \
  }                                  \
  catch (const DiffError& ex)        \
  {                                  \
    d->err = ex.what();              \
    return ex.code;                  \
  }                                  \
  catch (const std::exception& ex)   \
  {                                  \
    d->err = ex.what();              \
    return ESP_ERR_STATE;            \
  }

extern "C" {

const char* esp_diff_create_error(void) { return g_diff_create_err.c_str(); }
const char* esp_diff_last_error(const esp_diff* d) { return d ? d->err.c_str() : ""; }

int esp_diff_create(const esp_diff_config* cfg, esp_diff** out)
{
  g_diff_create_err.clear();
  if (!cfg || !out) {
    g_diff_create_err = "null argument";
    return ESP_ERR_ARG;
  }
  if (!(cfg->grid_spacing > 0) || !(cfg->verlet_dist >= 0) || !(cfg->dt > 0)) {
    g_diff_create_err = "gridSpacing, dt must be > 0 and verletDist >= 0";
    return ESP_ERR_ARG;
  }
  const int bsz = cfg->block > 0 ? cfg->block : 256;
  if (bsz < 64 || bsz > DIFF_MAX_BS || bsz % 64) {
    g_diff_create_err = "block must be a multiple of 64 in [64, 512]";
    return ESP_ERR_ARG;
  }
  // the neighbour table of TSubLattice::initNeighborTable (SubLattice.hpp:455-478)
  const double nr = cfg->grid_spacing;
  double nmin[3], nmax[3];
  long ng[3];
  for (int c = 0; c < 3; c++) {
    const double mn = cfg->domain_min[c], mx = cfg->domain_max[c];
    const double sz = nr * std::ceil((mx - mn) / nr);
    const double grow = sz - (mx - mn);
    nmin[c] = mn - grow * 0.5;
    nmax[c] = mx + grow * 0.5;
    ng[c] = lrint((nmax[c] - nmin[c]) / nr);
    if (std::fabs((double)ng[c] - (nmax[c] - nmin[c]) / nr) > 1e-6 || ng[c] < 1) {
      g_diff_create_err = "size doesn't fit range";
      return ESP_ERR_ARG;
    }
  }
  if ((double)(ng[0] + 2) * (double)(ng[1] + 2) * (double)(ng[2] + 2) > 2.0e8) {
    g_diff_create_err = "neighbour table too large";
    return ESP_ERR_CAPACITY;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_diff_create_err = "no HIP device available";
    return ESP_ERR_DEVICE;
  }
  if (cfg->device < 0 || cfg->device >= ndev || hipSetDevice(cfg->device) != hipSuccess) {
    g_diff_create_err = "bad device";
    return ESP_ERR_DEVICE;
  }
  esp_diff* d = new esp_diff();
  d->cfg = *cfg;
  d->bs = bsz;
  memset(&d->grid, 0, sizeof d->grid);
  d->grid.x0 = nmin[0]; d->grid.y0 = nmin[1]; d->grid.z0 = nmin[2];
  d->grid.dim = nr;
  d->grid.gx = d->grid.gy = d->grid.gz = -1;
  d->grid.nx = (int)ng[0] + 2; d->grid.ny = (int)ng[1] + 2; d->grid.nz = (int)ng[2] + 2;
  d->ncell = d->grid.nx * d->grid.ny * d->grid.nz;
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    g_diff_create_err = "hipStreamCreate failed";
    return ESP_ERR_DEVICE;
  }
  try {
    d->md.reserve(2);
    d->cell_count.reserve(d->ncell + 1);
    d->cell_start.reserve(d->ncell + 1);
    d->cell_fill.reserve(d->ncell + 1);
  } catch (const DiffError& ex) {
    g_diff_create_err = ex.what();
    delete d;
    return ex.code;
  }
  *out = d;
  return ESP_OK;
}

void esp_diff_destroy(esp_diff* d)
{
  if (!d) return;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  drop_graph(d);
  d->epochs.clear();
  d->ckpts.clear();
  hipStream_t s = d->stream;
  delete d;
  if (s) (void)hipStreamDestroy(s);
}

int esp_diff_set_particles(esp_diff* d, int n, int nspec, const int* spec, const double* pos,
                           const double* rad, const double* mass, const double* vel)
{
  DAPI_BEGIN
  if (n <= 0 || nspec <= 0 || !spec || !pos || !rad || !mass)
    throw DiffError(ESP_ERR_ARG, "bad particle arrays");
  std::vector<int> start(nspec + 1, 0);
  for (int i = 0; i < n; i++) {
    if (spec[i] < 0 || spec[i] >= nspec || (i > 0 && spec[i] < spec[i - 1]))
      throw DiffError(ESP_ERR_ARG, "specimen ids must be non-decreasing in [0, nspec)");
    start[spec[i] + 1]++;
  }
  for (int s = 0; s < nspec; s++) start[s + 1] += start[s];
  d->n = n;
  d->nspec = nspec;
  d->h_spec.assign(spec, spec + n);
  d->h_spec_start = start;
  d->h_pos.assign(pos, pos + 3 * n);
  d->h_rad.assign(rad, rad + n);
  d->h_mass.assign(mass, mass + n);
  d->h_vel.assign(vel ? vel : pos, (vel ? vel : pos) + 3 * n);
  if (!vel) std::fill(d->h_vel.begin(), d->h_vel.end(), 0.0);
  d->spec.reserve(n);
  d->spec_start.reserve(nspec + 1);
  d->id.reserve(n);
  d->flags.reserve(n);
  d->cell.reserve(n);
  d->cell_items.reserve(n);
  d->ccnt.reserve(n + 1);
  d->coff.reserve(n + 1);
  d->wfn.reserve(n);
  std::vector<int> ids(n);
  std::vector<unsigned> fl(n, PF_DYN | PF_ROT);
  for (int i = 0; i < n; i++) ids[i] = i;
  DCHK(hipMemcpyAsync(d->spec.p, spec, sizeof(int) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->spec_start.p, start.data(), sizeof(int) * (nspec + 1),
                      hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->id.p, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->flags.p, fl.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice,
                      d->stream));
  dsync(d);
  d->have_forward = false;
  return ESP_OK;
  DAPI_END
}

int esp_diff_set_bonds(esp_diff* d, int nb, const int* bi, const int* bj, const double* rel,
                       const double* tk, const double* bk)
{
  DAPI_BEGIN
  if (d->n <= 0) throw DiffError(ESP_ERR_STATE, "set the particles first");
  if (nb < 0 || (nb > 0 && (!bi || !bj || !rel)) || !tk || !bk)
    throw DiffError(ESP_ERR_ARG, "bad bond arrays");
  std::vector<int> order(nb);
  for (int k = 0; k < nb; k++) {
    if (bi[k] < 0 || bj[k] >= d->n || bi[k] >= bj[k] || d->h_spec[bi[k]] != d->h_spec[bj[k]])
      throw DiffError(ESP_ERR_ARG, "a bond needs i < j within one specimen");
    order[k] = k;
  }
  auto key = [&](int k) { return ((long long)bi[k] << 32) | (long long)bj[k]; };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int k = 1; k < nb; k++)
    if (key(order[k]) == key(order[k - 1])) throw DiffError(ESP_ERR_ARG, "repeated bond");
  std::vector<int2> ij(nb);
  std::vector<long long> kk(nb);
  std::vector<double4> D0(nb), R(nb);
  for (int t = 0; t < nb; t++) {
    const int k = order[t], a = bi[k], b = bj[k];
    ij[t] = make_int2(a, b);
    kk[t] = key(k);
    const double dx = d->h_pos[3 * b] - d->h_pos[3 * a], dy = d->h_pos[3 * b + 1] - d->h_pos[3 * a + 1],
                 dz = d->h_pos[3 * b + 2] - d->h_pos[3 * a + 2];
    // ARotBondedInteraction / CRotBondedInteraction: D = p2 - p1, dist = |D|
    D0[t] = make_double4(dx, dy, dz, std::sqrt(dx * dx + dy * dy + dz * dz));
    R[t] = make_double4(rel[4 * k], rel[4 * k + 1], rel[4 * k + 2], rel[4 * k + 3]);
  }
  std::vector<double2> tkb(d->nspec);
  for (int s = 0; s < d->nspec; s++) tkb[s] = make_double2(tk[s], bk[s]);
  d->nb = nb;
  const size_t c = std::max(nb, 1);
  d->bij.reserve(c); d->bkey.reserve(c); d->bD0.reserve(c); d->brel.reserve(c); d->bstep.reserve(c);
  d->tkb.reserve(std::max(d->nspec, 1));
  DCHK(hipMemcpyAsync(d->bij.p, ij.data(), sizeof(int2) * nb, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->bkey.p, kk.data(), sizeof(long long) * nb, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->bD0.p, D0.data(), sizeof(double4) * nb, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->brel.p, R.data(), sizeof(double4) * nb, hipMemcpyHostToDevice, d->stream));
  DCHK(hipMemcpyAsync(d->tkb.p, tkb.data(), sizeof(double2) * d->nspec, hipMemcpyHostToDevice,
                      d->stream));
  d->h_bi.resize(nb);
  d->h_bj.resize(nb);
  for (int t = 0; t < nb; t++) { d->h_bi[t] = ij[t].x; d->h_bj[t] = ij[t].y; }
  dsync(d);
  d->have_forward = false;
  return ESP_OK;
  DAPI_END
}

int esp_diff_add_wall(esp_diff* d, const double origin[3], const double normal[3], double k,
                      const double step[3], int measure, const double* area)
{
  DAPI_BEGIN
  if (d->nw >= DIFF_MAX_WALLS) throw DiffError(ESP_ERR_CAPACITY, "more than 4 walls");
  if (!origin || !normal || !step) throw DiffError(ESP_ERR_ARG, "bad wall");
  DWall& W = d->wall0[d->nw];
  for (int c = 0; c < 3; c++) {
    W.o[c] = origin[c];
    W.n[c] = normal[c];
    d->wstep[d->nw][c] = step[c];
  }
  W.k = k;
  W.measure = measure ? 1 : 0;
  if (measure) {
    if (!area || d->nspec <= 0) throw DiffError(ESP_ERR_ARG, "the measured wall needs the areas");
    for (int w = 0; w < d->nw; w++)
      if (d->wall0[w].measure) throw DiffError(ESP_ERR_ARG, "one measured wall");
    d->h_area.assign(area, area + d->nspec);
  }
  d->nw++;
  return ESP_OK;
  DAPI_END
}

int esp_diff_set_gravity(esp_diff* d, const double g[3])
{
  DAPI_BEGIN
  d->has_g = 1;
  for (int c = 0; c < 3; c++) d->g[c] = g[c];
  return ESP_OK;
  DAPI_END
}

int esp_diff_set_run(esp_diff* d, int nsteps, int sample_every)
{
  DAPI_BEGIN
  if (nsteps <= 0 || sample_every <= 0) throw DiffError(ESP_ERR_ARG, "bad run");
  d->nsteps = nsteps;
  d->every = sample_every;
  return ESP_OK;
  DAPI_END
}

int esp_diff_forward(esp_diff* d, const double* theta, const double* target, double* stress,
                     double* loss)
{
  DAPI_BEGIN
  if (!theta || !target || !loss) throw DiffError(ESP_ERR_ARG, "null argument");
  if (d->h_area.empty()) throw DiffError(ESP_ERR_STATE, "no measured wall");
  forward(d, theta, target, stress, loss);
  return ESP_OK;
  DAPI_END
}

int esp_diff_forward_tangent(esp_diff* d, const double* theta, const double* dtheta,
                             const double* target, double* stress, double* loss, double* dloss)
{
  DAPI_BEGIN
  if (!theta || !dtheta || !target || !loss || !dloss) throw DiffError(ESP_ERR_ARG, "null argument");
  if (d->h_area.empty()) throw DiffError(ESP_ERR_STATE, "no measured wall");
  forward(d, theta, target, stress, loss, dtheta, dloss);
  return ESP_OK;
  DAPI_END
}

int esp_diff_backward(esp_diff* d, double* grad)
{
  DAPI_BEGIN
  if (!grad) throw DiffError(ESP_ERR_ARG, "null argument");
  backward(d, grad);
  return ESP_OK;
  DAPI_END
}

int esp_diff_get_particle_data(esp_diff* d, int field, double* out)
{
  DAPI_BEGIN
  const int n = d->n;
  std::vector<double4> a(n);
  const double4* src = field == ESP_FIELD_POS ? d->x.p : field == ESP_FIELD_VEL ? d->v.p
                     : field == ESP_FIELD_ANGVEL ? d->w.p : field == ESP_FIELD_QUAT ? d->q.p : nullptr;
  if (!src) throw DiffError(ESP_ERR_ARG, "unknown field");
  DCHK(hipMemcpyAsync(a.data(), src, sizeof(double4) * n, hipMemcpyDeviceToHost, d->stream));
  dsync(d);
  for (int i = 0; i < n; i++) {
    if (field == ESP_FIELD_QUAT) {
      out[4 * i] = a[i].x; out[4 * i + 1] = a[i].y; out[4 * i + 2] = a[i].z; out[4 * i + 3] = a[i].w;
    } else {
      out[3 * i] = a[i].x; out[3 * i + 1] = a[i].y; out[3 * i + 2] = a[i].z;
    }
  }
  return ESP_OK;
  DAPI_END
}

int esp_diff_get_bond_breaks(esp_diff* d, int* step, int* bi, int* bj)
{
  DAPI_BEGIN
  std::vector<int> bs(std::max(d->nb, 1));
  DCHK(hipMemcpyAsync(bs.data(), d->bstep.p, sizeof(int) * d->nb, hipMemcpyDeviceToHost, d->stream));
  dsync(d);
  std::vector<int> idx;
  for (int b = 0; b < d->nb; b++)
    if (bs[b] != INT_MAX && bs[b] < d->nsteps) idx.push_back(b);  // left the group at bs + 1
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return bs[a] < bs[b]; });
  if (step)
    for (size_t k = 0; k < idx.size(); k++) {
      step[k] = bs[idx[k]] + 1;
      bi[k] = d->h_bi[idx[k]];
      bj[k] = d->h_bj[idx[k]];
    }
  return (int)idx.size();
  DAPI_END
}

int esp_diff_get_rebuild_steps(esp_diff* d, int* steps)
{
  DAPI_BEGIN
  if (steps)
    for (size_t k = 0; k < d->rebuild_steps.size(); k++) steps[k] = d->rebuild_steps[k];
  return (int)d->rebuild_steps.size();
  DAPI_END
}

int esp_diff_track_branches(esp_diff* d, int enable)
{
  DAPI_BEGIN
  d->track_digest = enable != 0;
  return ESP_OK;
  DAPI_END
}

int esp_diff_get_branch_digest(esp_diff* d, unsigned long long* out)
{
  DAPI_BEGIN
  if (!d->track_digest)
    throw DiffError(ESP_ERR_STATE, "branch digests are not tracked (esp_diff_track_branches)");
  if (!d->have_forward) throw DiffError(ESP_ERR_STATE, "no forward run");
  if (out) {
    DCHK(hipMemcpyAsync(out, d->digest.p + 1, sizeof(unsigned long long) * d->nsteps,
                        hipMemcpyDeviceToHost, d->stream));
    dsync(d);
  }
  return d->nsteps;
  DAPI_END
}

int esp_diff_get_times(esp_diff* d, double* fwd, double* bwd)
{
  DAPI_BEGIN
  if (fwd) *fwd = d->t_fwd;
  if (bwd) *bwd = d->t_bwd;
  return ESP_OK;
  DAPI_END
}

}  // extern "C"
