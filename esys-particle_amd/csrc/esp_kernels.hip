// esp_kernels.hip - HIP kernels of the per-timestep DEM engine (gfx950, FP64).
//
// Per step (launched back to back on one stream, no host synchronisation):
//   k_forces
//       the force laws of every interaction group in one launch (one block
//       segment per group), one thread per interaction; each writes its share
//       into both particles' slots.  In a queued step block 0 meanwhile applies
//       the previous step's bond breaks and appends the new pairs of the
//       flagged particles (and computes their forces)
//   k_gather_integrate
//       one thread per particle: wall and body forces, then the records of its
//       interactions in a fixed order (CSR), then the CParticle/CRotParticle
//       first-order update and the max-displacement reduction; per-block
//       partials of the wall forces
//   k_wall_finalize (once per queued batch)
//       fixed-order sums of the wall-force partials of every step of the batch
// Every per-step kernel first evaluates the previous step's status; if that
// step displaced a particle by more than verletDist/2 or broke a bond, the host
// must do bookkeeping first, so the kernel records the halt and returns, and
// all later queued steps return too.
//
// Rebuild / update kernels (launched from the host's bookkeeping path):
//   cell binning (counting sort + per-cell ordering), candidate pair search over
//   the 27-cell neighbourhood emitting pairs sorted by (id1, id2), interaction
//   storage updates (history carried over by key lookup), bond removal,
//   wall candidate refresh and the per-particle CSR of force records.
#include <cstring>
#include <hip/hip_ext.h>

#include "esp_internal.h"
#include "esp_math.h"

#define BS 256

static inline int nblk(long long n, int bs = BS) { return (int)((n + bs - 1) / bs); }

__device__ __forceinline__ dv3 xyz(double4 a) { return mk3(a.x, a.y, a.z); }

__device__ __forceinline__ bool step_skip(const StepCtl& c)
{
  if (!c.guarded) return false;
  const int halted = *(volatile int*)c.halted;
  const unsigned long long md = c.prev->maxdisp2;  // both loads in flight together
  if (halted) return true;
  if (md > c.thresh) {  // neighbour table must be rebuilt: host bookkeeping
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *c.reason = 1;
      *c.halted = c.step_id;
    }
    return true;
  }
  return false;
}

// The image of particle 2 nearest to particle 1 in a circular x dimension:
// shifted by the period when the two are more than half a period apart, one
// addition, as a ghost copy carries its shifted position (CParticle::setCircular,
// Particle.cpp:336-342); period 0: not circular.
__device__ __forceinline__ double4 x_image(double4 P1, double4 P2, double period)
{
  if (period > 0.0) {
    const double d = P1.x - P2.x;
    if (d > 0.5 * period) P2.x = P2.x + period;
    else if (d < -0.5 * period) P2.x = P2.x - period;
  }
  return P2;
}

__device__ __forceinline__ long long pair_key(int a, int b)
{
  return (long long)(((unsigned long long)(unsigned)a << 32) | (unsigned long long)(unsigned)b);
}

// ---------------------------------------------------------------------------
// pair-key hash sets (linear probing; bounded probes)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned hs_slot(unsigned long long k, unsigned mask)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (unsigned)k & mask;
}
__device__ __forceinline__ int hs_find(HashRef h, long long key)
{
  const unsigned long long k = (unsigned long long)key;
  unsigned s = hs_slot(k, h.mask);
  for (unsigned probe = 0; probe <= h.mask; probe++) {
    const unsigned long long v = h.t[s];
    if (v == k) return (int)s;
    if (v == HS_EMPTY) return -1;
    s = (s + 1) & h.mask;
  }
  return -1;
}
__device__ __forceinline__ void hs_insert(HashRef h, long long key)
{
  const unsigned long long k = (unsigned long long)key;
  unsigned s = hs_slot(k, h.mask);
  for (unsigned probe = 0; probe <= h.mask; probe++) {
    const unsigned long long old = atomicCAS(&h.t[s], HS_EMPTY, k);
    if (old == HS_EMPTY || old == k) return;
    s = (s + 1) & h.mask;
  }
}
__device__ __forceinline__ void hs_erase(HashRef h, long long key)
{
  const int s = hs_find(h, key);
  if (s >= 0) h.t[s] = HS_TOMB;
}
__device__ __forceinline__ bool excluded(const ExclSet& ex, long long key)
{
  for (int x = 0; x < ex.count; x++)
    if (hs_find(ex.h[x], key) >= 0) return true;
  return false;
}

// a candidate pair (particle slots a, b; id key) the group does not take
__device__ __forceinline__ bool rejected(const ExclSet& ex, long long key, int a, int b)
{
  if (ex.tagged) {
    int t1 = ex.tag[a], t2 = ex.tag[b];
    if (t1 > t2) {
      const int th = t1;
      t1 = t2;
      t2 = th;
    }
    if (!(((t1 & ex.mask1) == (ex.tag1 & ex.mask1)) && ((t2 & ex.mask2) == (ex.tag2 & ex.mask2))))
      return true;
  }
  return excluded(ex, key);
}

// ---------------------------------------------------------------------------
// per-step interaction kernels
// ---------------------------------------------------------------------------
// Each interaction writes its result per side straight into the two
// particles' CSR slots: force and moment on that particle (and, with stresses
// tracked, the contact point relative to it).  Slot positions are fixed at each
// rebuild (slot_of), so a particle later reads one contiguous run.
__device__ __forceinline__ void put3(double* d, dv3 v)
{
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
}
__device__ __forceinline__ void zero_slot(double* d, int ss)
{
  for (int k = 0; k < ss; k++) d[k] = 0.0;
}
// contact point p2 + r2 / (r1 + r2) (p1 - p2) of the pair laws
// (ARotBondedInteraction.cpp:437-440, RotFricInteraction.cpp:185, ElasticInteraction.cpp:61),
// stored relative to each particle
__device__ __forceinline__ void put_rel(double* a, double* b, double4 P1, double4 P2)
{
  const dv3 D = sub3(xyz(P1), xyz(P2));
  const dv3 c = add3(xyz(P2), mul3(D, P2.w / (P1.w + P2.w)));
  put3(a + 6, sub3(c, xyz(P1)));
  put3(b + 6, sub3(c, xyz(P2)));
}

// A bond's particle pair read with one 8-byte load.  The incremental update
// that runs beside the force blocks (block 0 of a queued step) marks the bonds
// broken in the previous step dead by storing (-1, -1) over their pairs.  A
// plain int2 read may be compiled into two 4-byte loads, the second issued
// only after the first index passed its test, and could then return a live
// first index with the dead second one (and the force law would gather
// position -1).  One relaxed atomic load returns either the whole old pair or
// (-1, -1).
__device__ __forceinline__ int2 load_pair_once(const int2* a)
{
  const unsigned long long v =
      __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<int2*>(a)),
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return make_int2((int)(unsigned)(v & 0xffffffffull), (int)(unsigned)(v >> 32));
}

// ARotBondedInteraction::calcForces + CRotBondedInteraction::broken
// (ARotBondedInteraction.cpp:307-448, RotBondedInteraction.cpp:175-219)
__device__ __forceinline__ void rot_bond_item(const StepCtl& ctl, const PView& p, int b,
                                              const ForceSeg& g, const ForceLaunch& fl)
{
  const int nb = g.n;
  const int2* __restrict__ ij = g.ij;
  if (b >= nb) return;
  const int2 e = load_pair_once(ij + b);
  const int2 so = fl.slot_of[g.rec_off + b];
  if (e.x < 0 || e.y < 0) return;  // broken earlier: removed from the group, slots left at zero
  const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], fl.period);
  const double4 Q1 = p.quat[e.x], Q2 = p.quat[e.y];
  const double4 K = g.k4[b], D = g.D4[b];
  // broken in an earlier step (D.w holds that step): the incremental update
  // that runs beside this launch removes the bond and clears its slots
  if (D.w != 0.0 && D.w < fl.step) return;
  BondResult o;
  esp_rot_bond_law(xyz(P1), dquat{Q1.x, mk3(Q1.y, Q1.z, Q1.w)}, P1.w, xyz(P2),
                   dquat{Q2.x, mk3(Q2.y, Q2.z, Q2.w)}, P2.w, K.x, K.y, K.z, K.w, xyz(D), o);
  // the strengths are needed only by the break test after the law: loaded
  // here, their latency overlaps the slot stores
  const double4 M = g.mx4[b];
  double* s0 = fl.slots + (size_t)fl.ss * so.x;
  double* s1 = fl.slots + (size_t)fl.ss * so.y;
  put3(s0, o.f1);
  put3(s0 + 3, o.m1);
  put3(s1, neg3(o.f1));
  put3(s1 + 3, o.m2);
  if (fl.ss == 9) put_rel(s0, s1, P1, P2);
  fl.sact[so.x] = 1;
  fl.sact[so.y] = 1;
  // CRotBondedInteraction::getCriterion (RotBondedInteraction.cpp:221-227)
  if (fl.crit)
    fl.crit[g.bond_off + b] = o.nForce / M.x + o.shForce / M.y + o.tMoment / M.z + o.bMoment / M.w;
  if (g.sc) return;  // broken by the particle stresses after the integration
  if (esp_bond_broken(o.nForce, o.shForce, o.tMoment, o.bMoment, M.x, M.y, M.z, M.w, g.truncated,
                      g.beta1, g.beta2)) {
    g.D4[b].w = fl.step;
    const int slot = atomicAdd(&ctl.cur->nbroken, 1);
    fl.broken_list[slot] = g.bond_off + b;
  }
}

__device__ __forceinline__ int record_of(int c, int nc, int rec_off, int app_rec_off)
{
  return c < nc ? rec_off + c : app_rec_off + (c - nc);
}

// CRotFrictionInteraction::calcForces (RotFricInteraction.cpp:103-231)
// RIGID: the launch may hold rigid=True groups (their law is only compiled into
// the kernel that handles them, so the common kernel keeps its registers)
template <bool RIGID>
__device__ __forceinline__ void rot_friction_item(const PView& p, int c, const ForceSeg& g,
                                                  const ForceLaunch& fl)
{
  const int nc = g.n;
  const int2 e = g.ij[c];
  const int rec = record_of(c, nc, g.rec_off, g.app_rec_off);
  const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], fl.period);
  const double4 st = g.ff[c];
  const dv3 Dq = sub3(xyz(P1), xyz(P2));
  const double eq = P1.w + P2.w;
  if (!(dot3(Dq, Dq) < (eq * eq))) {
    // not touching: history cleared, no force.  A pair that did not touch in
    // the last step already holds zero slots; once both history buffers hold
    // zero (flag bit 2) nothing is written (nor are its slots looked up).
    const int fw = (int)st.w;
    if ((fw & 3) != 0 || st.x != 0.0 || st.y != 0.0 || st.z != 0.0) {
      const int2 so = fl.slot_of[rec];
      g.ff_out[c] = make_double4(0.0, 0.0, 0.0, 0.0);
      zero_slot(fl.slots + (size_t)fl.ss * so.x, fl.ss);
      zero_slot(fl.slots + (size_t)fl.ss * so.y, fl.ss);
      fl.sact[so.x] = 0;
      fl.sact[so.y] = 0;
    } else if (!(fw & 4)) {
      g.ff[c] = make_double4(0.0, 0.0, 0.0, 4.0);
      g.ff_out[c] = make_double4(0.0, 0.0, 0.0, 4.0);
    }
    return;
  }
  // touching: the slot lookup goes out with the velocity loads
  const int2 so = fl.slot_of[rec];
  double* s0 = fl.slots + (size_t)fl.ss * so.x;
  double* s1 = fl.slots + (size_t)fl.ss * so.y;
  const double4 V1 = p.vel[e.x], V2 = p.vel[e.y], W1 = p.ang[e.x], W2 = p.ang[e.y];
  double kk, kks;
  esp_friction_k(g.k, g.ks, g.scaling, g.meanR, P1.w, P2.w, fl.do2d, kk, kks);
  dv3 ffr = mk3(st.x, st.y, st.z);
  bool slipping = (((int)st.w) & 1) != 0;
  dv3 f1, m1, m2;
  if (RIGID && g.rigid)
    esp_rot_friction_rigid_law(xyz(P1), xyz(V1), xyz(W1), P1.w, xyz(P2), xyz(V2), xyz(W2), P2.w,
                               kk, kks, g.mu_d, g.mu_s, fl.dt, (((int)st.w) & 2) != 0, ffr,
                               slipping, f1, m1, m2);
  else
    esp_rot_friction_law(xyz(P1), xyz(V1), xyz(W1), P1.w, xyz(P2), xyz(V2), xyz(W2), P2.w, kk,
                         kks, g.mu_d, g.mu_s, fl.dt, ffr, slipping, f1, m1, m2);
  g.ff_out[c] = make_double4(ffr.x, ffr.y, ffr.z, (double)((slipping ? 1 : 0) | 2));
  put3(s0, f1);
  put3(s0 + 3, m1);
  put3(s1, neg3(f1));
  put3(s1 + 3, m2);
  if (fl.ss == 9) put_rel(s0, s1, P1, P2);
  fl.sact[so.x] = 1;
  fl.sact[so.y] = 1;
}

// CElasticInteraction::calcForces (ElasticInteraction.cpp:54-71); no moments
// (their slot parts stay zero)
__device__ __forceinline__ void nrot_elastic_item(const PView& p, int c, const ForceSeg& g,
                                                  const ForceLaunch& fl)
{
  const int nc = g.n;
  const int2 e = g.ij[c];
  const int2 so = fl.slot_of[record_of(c, nc, g.rec_off, g.app_rec_off)];
  const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], fl.period);
  double* s0 = fl.slots + (size_t)fl.ss * so.x;
  double* s1 = fl.slots + (size_t)fl.ss * so.y;
  const dv3 D = sub3(xyz(P1), xyz(P2));
  double dist = dot3(D, D);
  const double eq_dist = P1.w + P2.w;
  if (dist < (eq_dist * eq_dist)) {
    const double kk = esp_elastic_k(g.k, g.scaling, P1.w, P2.w, fl.do2d);
    dist = sqrt(dist);
    const dv3 force = mul3(D, kk * (dist - eq_dist) / dist);
    put3(s0, neg3(force));
    put3(s1, force);
    if (fl.ss == 9) put_rel(s0, s1, P1, P2);
    fl.sact[so.x] = 1;
    fl.sact[so.y] = 1;
  } else {
    put3(s0, mk3(0.0, 0.0, 0.0));
    put3(s1, mk3(0.0, 0.0, 0.0));
    fl.sact[so.x] = 0;
    fl.sact[so.y] = 0;
  }
}

// CRotElasticInteraction::calcForces (RotElasticInteraction.cpp:69-88)
__device__ __forceinline__ void rot_elastic_item(const PView& p, int c, const ForceSeg& g,
                                                 const ForceLaunch& fl)
{
  const int nc = g.n;
  const int2 e = g.ij[c];
  const int2 so = fl.slot_of[record_of(c, nc, g.rec_off, g.app_rec_off)];
  const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], fl.period);
  double* s0 = fl.slots + (size_t)fl.ss * so.x;
  double* s1 = fl.slots + (size_t)fl.ss * so.y;
  const dv3 rb = sub3(xyz(P1), xyz(P2));
  const double rbNorm = norm3(rb);
  const double r_0Norm = P1.w + P2.w;
  if (rbNorm < r_0Norm) {
    const dv3 delta_r = div3(mul3(rb, rbNorm - r_0Norm), rbNorm);
    const dv3 force = mul3(delta_r, esp_rotelastic_k(g.k, g.scaling, P1.w, P2.w, fl.do2d));
    put3(s0, neg3(force));
    put3(s1, force);
    if (fl.ss == 9) put_rel(s0, s1, P1, P2);
    fl.sact[so.x] = 1;
    fl.sact[so.y] = 1;
  } else {
    put3(s0, mk3(0.0, 0.0, 0.0));
    put3(s1, mk3(0.0, 0.0, 0.0));
    fl.sact[so.x] = 0;
    fl.sact[so.y] = 0;
  }
}

// NeighborTable::index (ntable.hpp:128-139): floor((x - x0)/dim) - global index,
// with values outside the int range (and NaN) mapped to "outside" as x86's
// conversion does.
__device__ __forceinline__ bool cell_coords(const GridParams& g, double x, double y, double z,
                                            int& ix, int& iy, int& iz)
{
  const double fx = floor((x - g.x0) / g.dim);
  const double fy = floor((y - g.y0) / g.dim);
  const double fz = floor((z - g.z0) / g.dim);
  if (!(fx >= -2147483648.0 && fx < 2147483648.0 && fy >= -2147483648.0 && fy < 2147483648.0 &&
        fz >= -2147483648.0 && fz < 2147483648.0))
    return false;
  ix = (int)fx - g.gx;
  iy = (int)fy - g.gy;
  iz = (int)fz - g.gz;
  // circular x: positions are wrapped into [x0, x0 + L) at each rebuild; a
  // rounding of (x - x0)/dim onto the end of the range stays in the last cell
  if (g.circ_x && x >= g.x0 && x < g.x0 + g.period) ix = ix < 0 ? 0 : (ix >= g.nx ? g.nx - 1 : ix);
  return true;
}

// CBondedWallInteraction::calcForces (BWallInteraction.hpp:35-52): the force on
// the particle, -(total displacement of the particle - that of the wall) k',
// k' = k pi r (3-D; the constant of BWallInteraction.hpp:25-31) or k (2-D)
__device__ __forceinline__ dv3 bonded_wall_force(const WallParam& W, dv3 x, dv3 x_init, double r,
                                                 int do2d)
{
  const double scale = do2d ? 1.0 : 3.1415926536 * r;
  const double kb = W.k * scale;
  const dv3 D = sub3(sub3(x, x_init), mk3(W.disp[0], W.disp[1], W.disp[2]));
  return mul3(mul3(D, kb), -1.0);
}

// Wall-group membership and inner flags of one particle (CEWallInteractionGroup::
// Update, EWallInteractionGroup.hpp:57-82, with isInInner, ntable.hpp:207-222).
__device__ __forceinline__ unsigned wall_membership(const GridParams& g, int nwg,
                                                    const WallParam* W, double4 P, unsigned fl)
{
  fl &= ~((0xFFu << PF_WALL_SHIFT) | (0xFFu << PF_INNER_SHIFT));
  int ix = 0, iy = 0, iz = 0;
  bool inner = false;
  if (cell_coords(g, P.x, P.y, P.z, ix, iy, iz)) {
    // circular x: no padding cells in x, every x is inner
    const bool xin = g.circ_x || ((ix > 0) && (ix < g.nx - 1));
    if (g.nz < 3) inner = xin && (iy > 0) && (iy < g.ny - 1);
    else inner = xin && (iy > 0) && (iy < g.ny - 1) && (iz > 0) && (iz < g.nz - 1);
  }
  for (int w = 0; w < nwg; w++) {
    const dv3 n = mk3(W[w].nrm[0], W[w].nrm[1], W[w].nrm[2]);
    const double dist = dot3(sub3(xyz(P), mk3(W[w].o[0], W[w].o[1], W[w].o[2])), n);
    if (fabs(dist) < g.dim) {
      fl |= 1u << (PF_WALL_SHIFT + w);
      if (inner) fl |= 1u << (PF_INNER_SHIFT + w);
    }
  }
  return fl;
}

// ---------------------------------------------------------------------------
// gather + integrate
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// Fixed-order sum of (a, b, c) over the block, valid in thread 0: xor butterfly
// inside each wave (every lane ends with the same value; a + b == b + a), then
// the waves in index order.  Every thread of the block must call it.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* lds)
{
  for (int o = 32; o > 0; o >>= 1) {
    a = a + __shfl_xor(a, o, 64);
    b = b + __shfl_xor(b, o, 64);
    c = c + __shfl_xor(c, o, 64);
  }
  __syncthreads();  // lds free from the previous call
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    lds[3 * w] = a;
    lds[3 * w + 1] = b;
    lds[3 * w + 2] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < BS / 64; w++) {
      a = a + lds[3 * w];
      b = b + lds[3 * w + 1];
      c = c + lds[3 * w + 2];
    }
}

// sigma(i,j) -= f_i rel_j (CParticle::applyForce, Particle.cpp:391-400)
__device__ __forceinline__ void stress_add(double (&sg)[3][3], dv3 f, dv3 rel)
{
  const double fv[3] = {f.x, f.y, f.z}, rv[3] = {rel.x, rel.y, rel.z};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sg[i][j] -= fv[i] * rv[j];
}

// STRESS: records carry contact points and every applied force also updates
// the particle's stress sum, in the order the forces are summed
// EXTRA: damping groups or bonded walls exist (kept out of the plain kernel)
template <bool STRESS, bool EXTRA>
__global__ __launch_bounds__(BS) void k_gather_integrate(StepCtl ctl, PView p, GatherParams gp)
{
  constexpr int GATHER_CHUNK = STRESS ? 2 : 4;
  constexpr int GATHER_CHUNK6 = 8;  // active slots fetched together (6-double slots)
  constexpr int GATHER_WIN = 28;    // slots per activity-mask window (8 words with any offset)
  static_assert(GATHER_WIN < 32 && GATHER_WIN % 4 == 0, "a window's mask fits 32 bits");
  double sg[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sg[i][j] = 0.0;
  if (step_skip(ctl)) return;
  __shared__ double lds3[3 * (BS / 64)];
  __shared__ unsigned long long mred[BS / 64];
  // pairs appended by this step's incremental update are covered by the force
  // blocks from the next step on
  if (gp.n_app && blockIdx.x == 0 && threadIdx.x == 0) gp.n_app[1] = gp.n_app[0];
  const int i = blockIdx.x * BS + threadIdx.x;
  const bool act = i < p.n;
  const int ii = act ? i : 0;
  // every load that does not depend on another is issued here, together
  const double4 P = p.pos[ii];
  const unsigned fl = p.flags[ii];
  const double4 Vv = p.vel[ii];
  const double4 O = p.old[ii];
  const int e0 = gp.inc_off[ii], e1 = gp.inc_off[ii + 1];
  const int xcnt = gp.xcnt[ii];
  double4 A = make_double4(0.0, 0.0, 0.0, 0.0), Q = A;
  if (gp.rot) {
    A = p.ang[ii];
    Q = p.quat[ii];
  }
  const int refresh = gp.nwg > 0 && ctl.guarded && ctl.prev->nbroken > 0;
  const dv3 x0 = xyz(P);
  dv3 x = x0;
  const double r = P.w;
  dv3 F = mk3(0.0, 0.0, 0.0), M = mk3(0.0, 0.0, 0.0);
  unsigned flw = fl;
  if (act && refresh) {
    // bonds broke in the previous step: the reference's updateInteractions
    // re-Updates the wall groups (EWallInteractionGroup.hpp:57-82) before forces
    flw = wall_membership(gp.grid, gp.nwg, gp.wall, P, fl);
    p.flags[i] = flw;
  }
  // wall interaction groups, name order (EWallInteraction.hpp:42-78); in a
  // bonded group the tagged candidates are bonded instead
  for (int w = 0; w < gp.nwg; w++) {
    if (!act || !((flw >> (PF_WALL_SHIFT + w)) & 1u)) continue;
    const WallParam& W = gp.wall[w];
    if (EXTRA && W.tagged && !((flw >> (PF_BTAG_SHIFT + w)) & 1u)) continue;
    if (EXTRA && W.bonded && ((flw >> (PF_BTAG_SHIFT + w)) & 1u)) {
      const dv3 fb = bonded_wall_force(W, x, xyz(gp.init[i]), r, gp.do2d);
      F = add3(F, fb);
      if (STRESS) stress_add(sg, fb, sub3(x, x));
      continue;
    }
    const dv3 n = mk3(W.nrm[0], W.nrm[1], W.nrm[2]);
    const double dist = dot3(sub3(x, mk3(W.o[0], W.o[1], W.o[2])), n);
    if (dist < r) {
      double f = 1.0;
      if (!gp.do2d) f *= 3.141592654 * r;
      const dv3 force = mul3(n, (f * W.k) * (r - dist));
      F = add3(F, force);
      if (STRESS) stress_add(sg, force, sub3(sub3(x, mul3(n, dist)), x));
    }
  }
  unsigned long long md = 0;
  if (act) {
    dv3 v = xyz(Vv);
    const double invm = Vv.w, mass = O.w;
    // body forces, name order (BodyForceGroup.hpp:41-65), applied at the centre
    for (int g = 0; g < gp.ngrav; g++) {
      const dv3 fg = mul3(mk3(gp.g[g][0], gp.g[g][1], gp.g[g][2]), mass);
      F = add3(F, fg);
      if (STRESS) stress_add(sg, fg, sub3(x, x));
    }
    // the particle's slots in summation order (its CSR run, then its appendix
    // slots), fetched GATHER_CHUNK at a time so their loads are in flight
    // together; the sum itself stays sequential.  Moments of laws without
    // moments are zeros, and a sum that starts at +0 never becomes -0, so
    // adding them changes nothing.
    const int nmain = e1 - e0;
    const int cnt = nmain + xcnt;
    const int ss = STRESS ? 9 : 6;
    if (!STRESS) {
      // A zero slot (a contact apart) is skipped: adding +0 changes no sum.
      // The activity bytes (0 or 1) of up to GATHER_WIN slots of the run come
      // in as aligned 32-bit words in one round of loads and become a bit
      // mask; the active slots then load GATHER_CHUNK6 at a time, in run order.
      // 6-double slots are 16-byte aligned: three 16-byte loads each.
      const unsigned* sw = reinterpret_cast<const unsigned*>(gp.sact);
      for (int wb = 0; wb < cnt; wb += GATHER_WIN) {
        unsigned mask = 0;
        int base = 0;  // slot index of mask bit 0
        if (wb < nmain) {
          const int s0 = e0 + wb, nw = min(GATHER_WIN, nmain - wb), sh = s0 & 3;
          unsigned wv[GATHER_WIN / 4 + 1];
#pragma unroll
          for (int k = 0; k <= GATHER_WIN / 4; k++) wv[k] = 4 * k < sh + nw ? sw[(s0 >> 2) + k] : 0u;
          unsigned long long m = 0;
#pragma unroll
          for (int k = 0; k <= GATHER_WIN / 4; k++) {
            const unsigned v = wv[k];
            const unsigned long long m4 = (v & 1u) | ((v >> 7) & 2u) | ((v >> 14) & 4u) | ((v >> 21) & 8u);
            m |= m4 << (4 * k);
          }
          mask = (unsigned)(m >> sh) & ((1u << nw) - 1u);  // nw <= GATHER_WIN < 32
          base = s0;
        } else {
          // the appendix slots follow the run (at most ESP_XSLOTS of them)
          for (int t = 0; t < xcnt; t++)
            if (gp.sact[gp.xbase + ESP_XSLOTS * i + t]) mask |= 1u << t;
          base = gp.xbase + ESP_XSLOTS * i;
        }
        while (mask) {
          double2 qa[GATHER_CHUNK6], qb[GATHER_CHUNK6], qc[GATHER_CHUNK6];
          int sl[GATHER_CHUNK6];
          bool on[GATHER_CHUNK6];
#pragma unroll
          for (int k = 0; k < GATHER_CHUNK6; k++) {
            on[k] = mask != 0;
            sl[k] = base + (on[k] ? __builtin_ctz(mask) : 0);
            mask &= mask - 1u;
          }
#pragma unroll
          for (int k = 0; k < GATHER_CHUNK6; k++) {
            if (on[k]) {
              const double2* q = reinterpret_cast<const double2*>(gp.slots + 6 * (size_t)sl[k]);
              qa[k] = q[0];
              qb[k] = q[1];
              qc[k] = q[2];
            }
          }
#pragma unroll
          for (int k = 0; k < GATHER_CHUNK6; k++) {
            if (!on[k]) continue;
            F = add3(F, mk3(qa[k].x, qa[k].y, qb[k].x));
            M = add3(M, mk3(qb[k].y, qc[k].x, qc[k].y));
          }
        }
        // the appendix starts its own window right after the run
        if (wb < nmain && wb + GATHER_WIN >= nmain) wb = nmain - GATHER_WIN;
      }
    }
    for (int b = 0; STRESS && b < cnt; b += GATHER_CHUNK) {
      dv3 f[GATHER_CHUNK], m[GATHER_CHUNK], rl[GATHER_CHUNK];
#pragma unroll
      for (int k = 0; k < GATHER_CHUNK; k++) {
        const int t = b + k;
        const int slot = t >= cnt ? 0 : t < nmain ? e0 + t : gp.xbase + ESP_XSLOTS * i + (t - nmain);
        const double* q = gp.slots + (size_t)ss * slot;
        f[k] = mk3(q[0], q[1], q[2]);
        m[k] = mk3(q[3], q[4], q[5]);
        if (STRESS) rl[k] = mk3(q[6], q[7], q[8]);
      }
#pragma unroll
      for (int k = 0; k < GATHER_CHUNK; k++) {
        if (b + k >= cnt) break;
        F = add3(F, f[k]);
        if (STRESS) stress_add(sg, f[k], rl[k]);
        M = add3(M, m[k]);
      }
    }
    const double dt = gp.dt;
    // damping groups, last (SubLattice.hpp:2143-2147), in their map order
    for (int d = 0; EXTRA && d < gp.ndamp; d++) {
      const int ty = gp.damp_type[d];
      const double visc = gp.damp_visc[d];
      if (ty == ESP_DAMPING || ty == ESP_ROT_DAMPING) {
        const bool rd = ty == ESP_ROT_DAMPING;
        const double mi = rd ? gp.inert[i] : mass;
        const dv3 vr = esp_damping_vrel(rd ? xyz(A) : v, rd ? M : F, rd ? A.w : invm, mi, visc,
                                        dt, gp.damp_maxit[d]);
        const dv3 fd = mul3(mul3(vr, -1.0 * visc), mi);
        if (rd) {
          M = add3(M, fd);
        } else {
          F = add3(F, fd);
          if (STRESS) stress_add(sg, fd, sub3(x, x));
        }
      } else {
        const bool rd = ty == ESP_ROT_LOCAL_DAMPING;
        const dv3 vv = rd ? xyz(A) : v, ff = rd ? M : F;
        const dv3 fd = mul3(mk3(esp_local_damping_c(ff.x, vv.x), esp_local_damping_c(ff.y, vv.y),
                                esp_local_damping_c(ff.z, vv.z)),
                            -1.0 * visc);
        if (rd) {
          M = add3(M, fd);
        } else {
          F = add3(F, fd);
          if (STRESS) stress_add(sg, fd, sub3(x, x));
        }
      }
    }
    if (gp.rot) {
      dv3 w = xyz(A);
      dquat q = dquat{Q.x, mk3(Q.y, Q.z, Q.w)};
      if (fl & PF_ROT) {
        if (gp.do2d) {
          F = mk3(F.x, F.y, 0);
          M = mk3(0, 0, M.z);
        }
        w = add3(w, mul3(M, dt * A.w));
        const dquat dq = qmul4(dquat{0, w}, q);
        const double h = dt / 2.0;
        q.s += h * dq.s;
        q.v = add3(q.v, mul3(dq.v, h));
      } else {
        w = mk3(0.0, 0.0, 0.0);
      }
      if (fl & PF_DYN) {
        if (gp.do2d) F = mk3(F.x, F.y, 0);
        v = add3(v, mul3(mul3(F, invm), dt));
        x = add3(x, mul3(v, dt));
      }
      const double module = sqrt(q.s * q.s + q.v.x * q.v.x + q.v.y * q.v.y + q.v.z * q.v.z);
      if (module != 0.0) {
        const double inverse = 1.0 / module;
        q.s = inverse * q.s;
        q.v = mul3(q.v, inverse);
      }
      p.ang[i] = make_double4(w.x, w.y, w.z, A.w);
      p.quat[i] = make_double4(q.s, q.v.x, q.v.y, q.v.z);
      if (gp.store_fm) p.mom[i] = make_double4(M.x, M.y, M.z, 0.0);
    } else if (fl & PF_DYN) {
      if (gp.do2d) F = mk3(F.x, F.y, 0);
      v = add3(v, mul3(mul3(F, invm), dt));
      x = add3(x, mul3(v, dt));
    }
    p.pos[i] = make_double4(x.x, x.y, x.z, r);
    p.vel[i] = make_double4(v.x, v.y, v.z, invm);
    if (gp.store_fm) p.frc[i] = make_double4(F.x, F.y, F.z, 0.0);
    if (STRESS) {
      double* o = gp.sig + 9 * (size_t)i;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) o[3 * a + b] = sg[a][b];
    }
    const dv3 d = sub3(x, xyz(O));
    md = (unsigned long long)__double_as_longlong(dot3(d, d));
  }
  // max displacement: wave, then block, then one atomic per block
  md = wave_max_u64(md);
  if ((threadIdx.x & 63) == 0) mred[threadIdx.x >> 6] = md;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = mred[0];
    for (int w = 1; w < BS / 64; w++) m = mred[w] > m ? mred[w] : m;
    // only "beyond verletDist/2" is ever read, so smaller maxima need no atomic
    if (m > ctl.thresh) atomicMax(&ctl.cur->maxdisp2, m);
  }
  // this block's partial of each wall force: the forces on inner particles,
  // recomputed from the same values as above, in a fixed-order block sum;
  // a block without wall candidates writes zeros
  const unsigned wall_bits = act ? (flw >> PF_WALL_SHIFT) & 0xFFu : 0u;
  const int any_wall = gp.nwg > 0 ? __syncthreads_or(wall_bits != 0u) : 0;
  for (int w = 0; w < gp.nwg; w++) {
    if (!any_wall) {
      if (threadIdx.x == 0) {
        double* q = gp.wall_part + ((size_t)blockIdx.x * gp.nwg + w) * 3;
        q[0] = 0.0; q[1] = 0.0; q[2] = 0.0;
      }
      continue;
    }
    double wx = 0.0, wy = 0.0, wz = 0.0;
    if (act && ((flw >> (PF_WALL_SHIFT + w)) & 1u) && ((flw >> (PF_INNER_SHIFT + w)) & 1u)) {
      const WallParam& W = gp.wall[w];
      const dv3 n = mk3(W.nrm[0], W.nrm[1], W.nrm[2]);
      const double dist = dot3(sub3(x0, mk3(W.o[0], W.o[1], W.o[2])), n);
      if (EXTRA && W.tagged && !((flw >> (PF_BTAG_SHIFT + w)) & 1u)) {
        // untagged candidates of a tagged elastic wall: no interaction
      } else if (EXTRA && W.bonded && ((flw >> (PF_BTAG_SHIFT + w)) & 1u)) {
        // CWall::addForce(force) subtracts the bonded force: the particle's share
        const dv3 fb = bonded_wall_force(W, x0, xyz(gp.init[i]), r, gp.do2d);
        wx = fb.x; wy = fb.y; wz = fb.z;
      } else if (dist < r) {
        double f = 1.0;
        if (!gp.do2d) f *= 3.141592654 * r;
        const dv3 force = mul3(n, (f * W.k) * (r - dist));
        wx = force.x; wy = force.y; wz = force.z;
      }
    }
    block_sum3(wx, wy, wz, lds3);
    if (threadIdx.x == 0) {
      double* q = gp.wall_part + ((size_t)blockIdx.x * gp.nwg + w) * 3;
      q[0] = wx; q[1] = wy; q[2] = wz;
    }
  }
}

// Wall forces of a run of steps: block s sums step s's per-block partials in
// block order.  Launched once per queued batch; slots of steps that were
// skipped hold stale values and are not read by the host.
__global__ __launch_bounds__(BS) void k_wall_finalize(const double* part, int nblocks, int nwg,
                                                      double* out)
{
  __shared__ double lds3[3 * (BS / 64)];
  const double* ps = part + (size_t)blockIdx.x * nblocks * nwg * 3;
  for (int w = 0; w < nwg; w++) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += BS) {
      const double* q = ps + ((size_t)k * nwg + w) * 3;
      a = a + q[0];
      b = b + q[1];
      c = c + q[2];
    }
    block_sum3(a, b, c, lds3);
    if (threadIdx.x == 0) {
      double* o = out + ((size_t)blockIdx.x * nwg + w) * 3;
      o[0] = a;
      o[1] = b;
      o[2] = c;
    }
  }
}

__global__ void k_clear_status(StepStatus* ring, int ring_size, long long first_step, int count,
                               int* halted)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) {
    StepStatus& s = ring[(first_step + t) % ring_size];
    s.maxdisp2 = 0ull;
    s.nbroken = 0;
    s.pad = 0;
  }
  if (t == 0) *halted = 0;
}

// ---------------------------------------------------------------------------
// neighbour table
// ---------------------------------------------------------------------------
__global__ void k_cells_count(PView p, GridParams g, int* cell, int* cell_count, int* n_out)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= p.n) return;
  const double4 P = p.pos[i];
  int ix, iy, iz, c = -1;
  if (cell_coords(g, P.x, P.y, P.z, ix, iy, iz) && ix >= 0 && ix < g.nx && iy >= 0 && iy < g.ny &&
      iz >= 0 && iz < g.nz)
    c = g.ny * g.nz * ix + g.nz * iy + iz;
  cell[i] = c;
  if (c < 0) atomicAdd(n_out, 1);
  else atomicAdd(&cell_count[c], 1);
  // resetDisplacement (Particle.h:124)
  const double4 O = p.old[i];
  p.old[i] = make_double4(P.x, P.y, P.z, O.w);
}

__global__ void k_cells_fill(int n, const int* cell, const int* cell_start, int* cell_fill,
                             int* cell_items)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const int c = cell[i];
  if (c < 0) return;
  const int k = atomicAdd(&cell_fill[c], 1);
  cell_items[cell_start[c] + k] = i;
}

// cell contents in particle-array order
__global__ void k_cells_sort(int ncell, const int* cell_start, int* items)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= ncell) return;
  const int a = cell_start[c], b = cell_start[c + 1];
  for (int i = a + 1; i < b; i++) {
    const int v = items[i];
    int j = i - 1;
    while (j >= a && items[j] > v) {
      items[j + 1] = items[j];
      j--;
    }
    items[j + 1] = v;
  }
}


// Pairs of particle p (taken in ascending id order) with every particle q of
// higher id in the 27 cells around p's cell: the full list keeps pairs with
// |x_p - x_q|^2 <= (r_p + r_q + alpha)^2 (ntable.hpp:471-521); the flagged list
// keeps pairs of two flagged particles without a distance test
// (ntable.hpp:532-583).  The emit pass sorts each particle's run by the partner
// id, so the whole list comes out sorted by (id1, id2).
__global__ void k_pairs(int emit, int flagged, PView p, GridParams g, double alpha,
                        const int* cell, const int* cell_start, const int* cell_items, int* cnt,
                        const int* off, int* out_p, int* out_q, long long* out_key)
{
  const int k = blockIdx.x * BS + threadIdx.x;
  if (k >= p.n) return;
  const int pi = k;
  if (flagged && !(p.flags[pi] & PF_FLAG)) {
    if (!emit) cnt[k] = 0;
    return;
  }
  const int c = cell[pi];
  const int cx = c / (g.ny * g.nz);
  const int rem = c - cx * (g.ny * g.nz);
  const int cy = rem / g.nz;
  const int cz = rem - cy * g.nz;
  const double4 P = p.pos[pi];
  const int idp = p.id[pi];
  const int base = emit ? off[k] : 0;
  int count = 0;
  // circular x: the stencil wraps (nx >= 3, so its cells are distinct)
  const int x0 = g.circ_x ? cx - 1 : (cx > 0 ? cx - 1 : cx);
  const int x1 = g.circ_x ? cx + 1 : (cx < g.nx - 1 ? cx + 1 : cx);
  const int y0 = cy > 0 ? cy - 1 : cy, y1 = cy < g.ny - 1 ? cy + 1 : cy;
  const int z0 = cz > 0 ? cz - 1 : cz, z1 = cz < g.nz - 1 ? cz + 1 : cz;
  for (int xx = x0; xx <= x1; xx++)
    for (int y = y0; y <= y1; y++)
      for (int z = z0; z <= z1; z++) {
        const int x = g.circ_x ? (xx + g.nx) % g.nx : xx;
        const int c2 = g.ny * g.nz * x + g.nz * y + z;
        const int t1 = cell_start[c2 + 1];
        for (int t = cell_start[c2]; t < t1; t++) {
          const int q = cell_items[t];
          const int idq = p.id[q];
          if (idq <= idp) continue;
          if (flagged) {
            if (!(p.flags[q] & PF_FLAG)) continue;
          } else {
            const double4 Q = x_image(P, p.pos[q], g.period);
            const dv3 d = sub3(xyz(P), xyz(Q));
            const double dist2 = dot3(d, d);
            const double dmax = P.w + Q.w + alpha;
            if (!(dist2 <= (dmax * dmax))) continue;
          }
          if (emit) {
            out_p[base + count] = pi;
            out_q[base + count] = q;
            out_key[base + count] = pair_key(pi, q);
          }
          count++;
        }
      }
  if (!emit) {
    cnt[k] = count;
    return;
  }
  for (int i = base + 1; i < base + count; i++) {
    const long long kv = out_key[i];
    const int qv = out_q[i];
    int j = i - 1;
    while (j >= base && out_key[j] > kv) {
      out_key[j + 1] = out_key[j];
      out_q[j + 1] = out_q[j];
      j--;
    }
    out_key[j + 1] = kv;
    out_q[j + 1] = qv;
  }
}

// ---------------------------------------------------------------------------
// erasure of particles outside the padded grid (NeighborTable::build,
// ntable.hpp:281-286): stable compaction in slot order
// ---------------------------------------------------------------------------
__global__ void k_keep_cells(int n, const int* cell, int* keep)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) keep[i] = cell[i] >= 0 ? 1 : 0;
}

// new slot of every particle, -1 for an erased one
__global__ void k_new_slot(int n, const int* keep, const int* off, int* nslot)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) nslot[i] = keep[i] ? off[i] : -1;
}

template <class T>
__global__ void k_compact_field(int n, int width, const int* keep, const int* off, const T* src,
                                T* dst)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || !keep[i]) return;
  for (int c = 0; c < width; c++) dst[(size_t)width * off[i] + c] = src[(size_t)width * i + c];
}

// an interaction's particles at their new slots; kept while both survive
// (PIS_E::rebuild, pi_storage_e.hpp:111-155)
__global__ void k_remap_pairs(int nc, const int2* ij, const int* nslot, int2* out, int* keep)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc) return;
  const int2 e = ij[c];
  const int a = e.x >= 0 ? nslot[e.x] : -1, b = e.y >= 0 ? nslot[e.y] : -1;
  out[c] = make_int2(a, b);
  keep[c] = (a >= 0 && b >= 0) ? 1 : 0;
}

// the kept pairs of a dynamic group with their slot keys (order kept: slots
// are renumbered monotonically, so the list stays sorted by key)
__global__ void k_pack_dyn(int nc, const int* keep, const int* off, const int2* ij,
                           const double4* ff, int2* oij, long long* okey, double4* off_)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc || !keep[c]) return;
  const int w = off[c];
  const int2 e = ij[c];
  oij[w] = e;
  okey[w] = pair_key(e.x, e.y);
  if (off_) off_[w] = ff[c];
}

// the candidate particles of wall group w (flag bit PF_WALL_SHIFT + w): slot,
// flags, position and (bonded walls) initial position, in any order; all are
// counted, the first cap are stored
__global__ void k_wall_cands(PView p, int w, const double4* init, int cap, int* cnt, int* slot,
                             unsigned* flg, double4* pos, double4* ipos)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= p.n) return;
  const unsigned f = p.flags[i];
  if (!((f >> (PF_WALL_SHIFT + w)) & 1u)) return;
  const int k = atomicAdd(cnt, 1);
  if (k >= cap) return;
  slot[k] = i;
  flg[k] = f;
  pos[k] = p.pos[i];
  if (init) ipos[k] = init[i];
}

__global__ void k_wall_mark_ordered(int no, const int* ord_slot, const unsigned* flags, int w,
                                    int* mark)
{
  const int o = blockIdx.x * BS + threadIdx.x;
  if (o >= no) return;
  const int s = ord_slot[o];
  mark[o] = (s >= 0 && ((flags[s] >> (PF_WALL_SHIFT + w)) & 1u)) ? 1 : 0;
}

__global__ void k_wall_pack_ordered(int no, const int* ord_slot, const int* mark, const int* off,
                                    PView p, const double4* init, WallCand* out)
{
  const int o = blockIdx.x * BS + threadIdx.x;
  if (o >= no || !mark[o]) return;
  const int s = ord_slot[o];
  const double4 x0 = init ? init[s] : make_double4(0.0, 0.0, 0.0, 0.0);
  WallCand& c = out[off[o]];
  c.pos = p.pos[s];
  c.ix = x0.x;
  c.iy = x0.y;
  c.iz = x0.z;
  c.flags = p.flags[s];
  c.pad = 0u;
}

__global__ void k_clear_flag_bit(PView p, UpdState* st)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < p.n) p.flags[i] &= ~PF_FLAG;
  if (i == 0) st->fcount = 0;
}

__device__ __forceinline__ int bsearch_key(const long long* a, int n, long long key)
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
// number of elements < key (strict) or <= key
__device__ __forceinline__ int lower_bound_key(const long long* a, int n, long long key)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int upper_bound_key(const long long* a, int n, long long key)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// PIS_ED::update on a rebuilt table (pi_storage_ed.hpp:97-136): a candidate is
// kept unless excluded; it inherits the history of the stored interaction with
// the same key if that one is persistent (RotFricInteraction.cpp:345-351).
__global__ void k_ed_mark(int nc, const int* cp, const int* cq, const long long* ckey, ExclSet ex,
                          const long long* okey, int nold, const int2* oij, PView p, double period,
                          int* keep, int* src)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc) return;
  const long long key = ckey[c];
  if (rejected(ex, pair_key(p.id[cp[c]], p.id[cq[c]]), cp[c], cq[c])) {
    keep[c] = 0;
    src[c] = -1;
    return;
  }
  keep[c] = 1;
  int s = -1;
  const int j = bsearch_key(okey, nold, key);
  if (j >= 0) {
    const int2 e = oij[j];
    const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], period);
    const dv3 D = sub3(xyz(P1), xyz(P2));
    const double dist = dot3(D, D);
    const double eq = P1.w + P2.w;
    if (dist <= (eq * eq)) s = j;
  }
  src[c] = s;
}

__global__ void k_ne_mark(int nc, const int* cp, const int* cq, const int* id, ExclSet ex, int* keep)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c < nc) keep[c] = rejected(ex, pair_key(id[cp[c]], id[cq[c]]), cp[c], cq[c]) ? 0 : 1;
}

__global__ void k_pair_emit(int nc, const int* keep, const int* off, const int* cp, const int* cq,
                            const long long* ckey, const int* src, const double4* off_ff,
                            int2* nij, long long* nkey, double4* nff)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc || !keep[c]) return;
  const int w = off[c];
  nij[w] = make_int2(cp[c], cq[c]);
  nkey[w] = ckey[c];
  if (nff) {
    const int s = src ? src[c] : -1;
    nff[w] = s >= 0 ? off_ff[s] : make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

// pairs of the flagged list that a group takes: not excluded and (for history
// groups) not already stored (pi_storage_ed.hpp:137-156, pi_storage_ne.hpp:79-96)
__global__ void k_filter_new(int nc, const int* cp, const int* cq, const int* id,
                             const long long* ckey, ExclSet ex, HashRef present, int check_present,
                             int* keep)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc) return;
  int k = rejected(ex, pair_key(id[cp[c]], id[cq[c]]), cp[c], cq[c]) ? 0 : 1;
  if (k && check_present && hs_find(present, ckey[c]) >= 0) k = 0;
  keep[c] = k;
}

__global__ void k_pack_pairs(int n, const int* keep, const int* off, const int* cp, const int* cq,
                             const long long* ckey, int2* ij, long long* key)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= n || !keep[c]) return;
  const int w = off[c];
  ij[w] = make_int2(cp[c], cq[c]);
  key[w] = ckey[c];
}

// merge of two key-sorted lists; equal keys keep stored entries first
__global__ void k_merge(int nold, const int2* oij, const long long* okey, const double4* off_,
                        int nnew, const int2* nij_in, const long long* nkey_in,
                        const double4* nff_in, int2* ij_out, long long* key_out, double4* ff_out)
{
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t < nold) {
    const long long key = okey[t];
    const int w = t + lower_bound_key(nkey_in, nnew, key);
    ij_out[w] = oij[t];
    key_out[w] = key;
    if (ff_out) ff_out[w] = off_[t];
  } else if (t < nold + nnew) {
    const int j = t - nold;
    const long long key = nkey_in[j];
    const int w = j + upper_bound_key(okey, nold, key);
    ij_out[w] = nij_in[j];
    key_out[w] = key;
    if (ff_out) ff_out[w] = nff_in ? nff_in[j] : make_double4(0.0, 0.0, 0.0, 0.0);
  }
}


__global__ void k_fill_int(int n, int* a, int v)
{
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t < n) a[t] = v;
}

__global__ void k_compact_bonds(int nb, const int* keep, const int* off, const int2* ij,
                                const double4* k4, const double4* mx4, const double4* D4,
                                const int* rank, int2* ij2, double4* k42, double4* mx42,
                                double4* D42, int* rank2)
{
  const int b = blockIdx.x * BS + threadIdx.x;
  if (b >= nb || !keep[b]) return;
  const int w = off[b];
  ij2[w] = ij[b];
  k42[w] = k4[b];
  mx42[w] = mx4[b];
  D42[w] = D4[b];
  rank2[w] = rank[b];
}



// CEWallInteractionGroup::Update (EWallInteractionGroup.hpp:57-82): particles
// with |(x - o).n| < gridSpacing become candidates; isInInner (ntable.hpp:207-222)
// decides whether their force counts towards the wall force.
// The walls travel as one by-value array in the kernel arguments and are read
// in place (a local copy of the array would live in scratch memory).
struct WallSet {
  WallParam w[ESP_MAX_WALL_GROUPS];
};
__global__ void k_wall_update(PView p, GridParams g, int nwg, WallSet ws)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= p.n) return;
  p.flags[i] = wall_membership(g, nwg, ws.w, p.pos[i], p.flags[i]);
}

// ---------------------------------------------------------------------------
// bookkeeping of bond breaks on the device
// ---------------------------------------------------------------------------
// PIS_EB::update (pi_storage_eb.hpp:23-49) for one bond whose criterion held in
// the previous step: the bond leaves its group (marked dead, record zeroed, key
// erased from the exclusion set), both particles are flagged
// (RotBondedInteraction.cpp:211-213) and the event is logged.
__device__ __forceinline__ void break_one(PView p, const BondSet& bs, int gb, UpdState* st,
                                          int* flist, int4* evt, long long step)
{
  int g = 0;
  while (g + 1 < bs.count && gb >= bs.off[g + 1]) g++;
  const int b = gb - bs.off[g];
  if (b < 0 || b >= bs.n[g]) return;
  const int2 e = bs.ij[g][b];
  if (e.x < 0) return;
  // independent memory operations first, so their latencies overlap
  const int id1 = p.id[e.x], id2 = p.id[e.y];
  const int2 so = bs.slot_of[bs.rec_off[g] + b];
  const unsigned ox = atomicOr(&p.flags[e.x], PF_FLAG);
  const unsigned oy = atomicOr(&p.flags[e.y], PF_FLAG);
  bs.ij[g][b] = make_int2(-1, -1);
  const int ev = atomicAdd(&st->evt_count, 1);
  const int nx = (ox & PF_FLAG) ? 0 : 1, ny = (oy & PF_FLAG) ? 0 : 1;
  if (nx + ny) {
    const int f0 = atomicAdd(&st->fcount, nx + ny);
    if (nx) flist[f0] = e.x;
    if (ny) flist[f0 + nx] = e.y;
  }
  double* s0 = bs.slots + (size_t)bs.ss * so.x;
  double* s1 = bs.slots + (size_t)bs.ss * so.y;
  for (int k = 0; k < bs.ss; k++) {
    s0[k] = 0.0;
    s1[k] = 0.0;
  }
  bs.sact[so.x] = 0;
  bs.sact[so.y] = 0;
  evt[ev] = make_int4((int)step, id1, id2, 0);
  hs_erase(bs.hs[g], pair_key(id1, id2));
}

// the previous step's breaks on the host bookkeeping path
__global__ void k_break_apply(StepCtl ctl, PView p, BondSet bs, const int* broken_list,
                              UpdState* st, int* flist, int4* evt, long long step)
{
  if (step_skip(ctl)) return;
  const int nb = ctl.prev->nbroken;
  for (int t = blockIdx.x * BS + threadIdx.x; t < nb; t += gridDim.x * BS)
    break_one(p, bs, broken_list[t], st, flist, evt, step);
}

// outside the table's padded grid: erased by the next table build (the
// predicate of k_cells_count)
__device__ __forceinline__ bool outside_grid(const GridParams& g, const double4& P)
{
  int ix, iy, iz;
  return !(cell_coords(g, P.x, P.y, P.z, ix, iy, iz) && ix >= 0 && ix < g.nx && iy >= 0 &&
           iy < g.ny && iz >= 0 && iz < g.nz);
}

// k_break_apply before a table build that erases particles: the reference
// erases the particles outside the grid, with their bonds, before it tests the
// bonds (NeighborTable::build, then PIS_EB::update), so a bond with such a
// particle is left to the erasure: no break event, no flags.
__global__ void k_break_apply_erase(StepCtl ctl, PView p, BondSet bs, const int* broken_list,
                                    UpdState* st, int* flist, int4* evt, long long step,
                                    GridParams g)
{
  if (step_skip(ctl)) return;
  const int nb = ctl.prev->nbroken;
  for (int t = blockIdx.x * BS + threadIdx.x; t < nb; t += gridDim.x * BS) {
    const int gb = broken_list[t];
    int k = 0;
    while (k + 1 < bs.count && gb >= bs.off[k + 1]) k++;
    const int b = gb - bs.off[k];
    if (b < 0 || b >= bs.n[k]) continue;
    const int2 e = bs.ij[k][b];
    if (e.x >= 0 && (outside_grid(g, p.pos[e.x]) || outside_grid(g, p.pos[e.y]))) continue;
    break_one(p, bs, gb, st, flist, evt, step);
  }
}

// TSubLattice::updateInteractions after bond breaks, for a queued step, run by
// block 0 of the step's force launch while the other blocks compute the stored
// interactions: the previous step's breaks are applied (break_one), then
// getNewList (ntable.hpp:628-664) and the first dynamic group's incremental
// update (pi_storage_ed.hpp:137-156 / pi_storage_ne.hpp:79-96): pairs of
// flagged particles in neighbouring cells of the last build, without a distance
// test, minus excluded (and, for history groups, stored) pairs, are appended in
// key order and their forces of this step are computed here; then every flag
// is cleared.  The force blocks cover the appendix as it was before this step
// (n_app[1]; the gather advances it to n_app[0]).  Bonds broken in the
// previous step carry that step in D4.w, so the force blocks skip them without
// waiting for this block.  When the pairs do not fit the appendix, the queue
// halts and the host redoes the step (its input history buffer is intact).
#define NEW_SENTINEL 0x7fffffffffffffffll
template <bool RIGID>
__device__ __forceinline__ void dyn_item(const PView& p, int c, const ForceSeg& g,
                                         const ForceLaunch& fl)
{
  switch (g.kind) {
    case ESP_SEG_ROT_FRICTION: rot_friction_item<RIGID>(p, c, g, fl); break;
    case ESP_SEG_NROT_ELASTIC: nrot_elastic_item(p, c, g, fl); break;
    case ESP_SEG_ROT_ELASTIC: rot_elastic_item(p, c, g, fl); break;
    default: break;
  }
}

// RIGID, TRACK: the extended launch (rigid=True RotFriction groups; app_step,
// the pair-event log's step of each appended pair)
template <bool RIGID, bool TRACK>
__device__ void incremental_update(const StepCtl& ctl, const PView& p, const UpdParams& up,
                                   const ForceLaunch& fl, int* app_step)
{
  __shared__ long long s_key[ESP_NEW_CAP];
  __shared__ int s_cnt, s_over, s_keep;
  const int nbk = ctl.prev->nbroken;
  if (nbk <= 0) return;
  const int tid = threadIdx.x;
  for (int t = tid; t < nbk; t += BS)
    break_one(p, up.bs, up.broken_prev[t], up.st, up.flist, up.evt, up.step);
  if (!up.has_group) return;
  const DynAppend& da = up.da;
  const GridParams& g = up.g;
  // flags, flagged list and its count complete: within one workgroup a barrier
  // orders global memory (the waves share one vector L1)
  __syncthreads();
  if (tid == 0) {
    s_cnt = 0;
    s_over = 0;
    s_keep = 0;
  }
  __syncthreads();
  const int nf = up.st->fcount;
  // one wave per flagged particle; lane l scans half (l / 27) of the items of
  // neighbour cell l % 27, so the dependent loads of a cell scan run in parallel
  const int lane = tid & 63;
  for (int f = tid >> 6; f < nf; f += BS / 64) {
    const int pi = up.flist[f];
    const int c = up.cell[pi];
    const int cx = c / (g.ny * g.nz);
    const int rem = c - cx * (g.ny * g.nz);
    const int cy = rem / g.nz;
    const int cz = rem - cy * g.nz;
    const int idp = p.id[pi];
    if (lane >= 54) continue;
    const int nc = lane % 27, half = lane / 27;
    int x = cx + nc / 9 - 1;
    const int y = cy + (nc / 3) % 3 - 1, z = cz + nc % 3 - 1;
    if (g.circ_x) x = (x + g.nx) % g.nx;  // the stencil wraps in a circular x
    if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz) continue;
    const int c2 = g.ny * g.nz * x + g.nz * y + z;
    const int t0 = up.cell_start[c2], t1 = up.cell_start[c2 + 1];
    // this lane's items u = t0 + half + 2 j, four at a time with their loads in flight together
    for (int u = t0 + half; u < t1; u += 8) {
      int q[4];
      unsigned fq[4];
      int idq[4];
#pragma unroll
      for (int j = 0; j < 4; j++) q[j] = u + 2 * j < t1 ? up.cell_items[u + 2 * j] : -1;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        fq[j] = q[j] >= 0 ? p.flags[q[j]] : 0u;
        idq[j] = q[j] >= 0 ? p.id[q[j]] : 0;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (!(fq[j] & PF_FLAG) || idq[j] <= idp) continue;
        const int k = atomicAdd(&s_cnt, 1);
        if (k < ESP_NEW_CAP) {
          s_key[k] = pair_key(pi, q[j]);
        } else {
          s_over = 1;
        }
      }
    }
  }
  __syncthreads();
  const int n = min(s_cnt, ESP_NEW_CAP);
  // keep: not excluded, not stored (history groups); others sort to the end.
  // A key holds the pair's slots (lower-id particle first).
  int m = 1;
  while (m < n) m <<= 1;
  for (int t = tid; t < m; t += BS) {
    if (t < n) {
      const long long key = s_key[t];
      const int a = (int)(key >> 32), b = (int)(key & 0xffffffffll);
      const bool keep = !rejected(da.ex, pair_key(p.id[a], p.id[b]), a, b) &&
                        !(da.is_hist && hs_find(da.present, key) >= 0);
      if (keep) atomicAdd(&s_keep, 1);
      else s_key[t] = NEW_SENTINEL;
    } else {
      s_key[t] = NEW_SENTINEL;
    }
  }
  __syncthreads();
  if (n <= BS) {
    // few candidates: each kept key moves to its rank among the kept keys (keys
    // are distinct, the sentinel is larger than any), one pass
    long long mine = NEW_SENTINEL;
    int rank = 0;
    if (tid < n) {
      mine = s_key[tid];
      if (mine != NEW_SENTINEL)
        for (int u = 0; u < n; u++) rank += s_key[u] < mine ? 1 : 0;
    }
    __syncthreads();
    if (mine != NEW_SENTINEL) s_key[rank] = mine;
    __syncthreads();
  } else {
    for (int k = 2; k <= m; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < m; i += BS) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const bool up_ = (i & k) == 0;
            const long long a = s_key[i], b = s_key[ixj];
            if ((a > b) == up_) {
              s_key[i] = b;
              s_key[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
  }
  const int kept = s_keep;
  const int napp = da.n_app[0];
  if (napp + kept > ESP_APP_CAP) s_over = 1;
  // each new record takes the next appendix slot of both its particles, pairs
  // in key order: the slot is the count of earlier kept pairs on that particle
  for (int t = tid; t < kept; t += BS) {
    const int a = (int)(s_key[t] >> 32), b = (int)(s_key[t] & 0xffffffffll);
    int da_ = 0, db_ = 0;
    for (int u = 0; u < kept; u++) {
      const long long ku = s_key[u];
      const int a2 = (int)(ku >> 32), b2 = (int)(ku & 0xffffffffll);
      da_ += (a2 == a) | (b2 == a);
      db_ += (a2 == b) | (b2 == b);
    }
    if (up.xcnt[a] + da_ > ESP_XSLOTS || up.xcnt[b] + db_ > ESP_XSLOTS) s_over = 1;
  }
  __syncthreads();
  if (s_over) {
    if (tid == 0) {
      *ctl.reason = 2;
      *ctl.halted = ctl.step_id;
    }
    return;
  }
  for (int t = tid; t < kept; t += BS) {
    const long long key = s_key[t];
    const int a = (int)(key >> 32), b = (int)(key & 0xffffffffll);
    int pa = 0, pb = 0;
    for (int u = 0; u < t; u++) {
      const long long ku = s_key[u];
      const int a2 = (int)(ku >> 32), b2 = (int)(ku & 0xffffffffll);
      pa += (a2 == a) | (b2 == a);
      pb += (a2 == b) | (b2 == b);
    }
    const int w = da.n_main + napp + t;
    da.ij[w] = make_int2(a, b);
    da.key[w] = key;
    if (da.ff) da.ff[w] = make_double4(0.0, 0.0, 0.0, 0.0);
    if (TRACK && app_step) app_step[napp + t] = (int)up.step;
    if (da.is_hist) hs_insert(da.present, key);
    // appendix slots start at zero (cleared at the last rebuild, used once)
    da.slot_of[da.app_rec_off + napp + t] =
        make_int2(da.xbase + ESP_XSLOTS * a + up.xcnt[a] + pa, da.xbase + ESP_XSLOTS * b + up.xcnt[b] + pb);
  }
  __syncthreads();  // every xcnt read above precedes the increments
  for (int t = tid; t < kept; t += BS) {
    const long long key = s_key[t];
    atomicAdd(&up.xcnt[(int)(key >> 32)], 1);
    atomicAdd(&up.xcnt[(int)(key & 0xffffffffll)], 1);
  }
  // this step's forces of the new pairs (each entry was written by the same thread)
  const ForceSeg& us = fl.seg[up.useg];
  const int c0 = da.n_main + napp;
  for (int t = tid; t < kept; t += BS) dyn_item<RIGID>(p, c0 + t, us, fl);
  if (tid == 0) da.n_app[0] = napp + kept;
  for (int t = tid; t < nf; t += BS) p.flags[up.flist[t]] &= ~PF_FLAG;
  if (tid == 0) up.st->fcount = 0;
}

// The force laws of every interaction group in one launch: each segment of
// blocks covers one group (bonded groups first, their law is the longest), so
// the groups' memory latencies overlap.  Every interaction writes its slots.
// In a queued step block 0 runs the incremental update (fl.ublk).
template <bool RIGID, bool TRACK>
__device__ __forceinline__ void forces_body(const StepCtl& ctl, const PView& p,
                                            const ForceLaunch& fl, const UpdParams& up,
                                            int* app_step)
{
  if (step_skip(ctl)) return;
  if (fl.ublk && blockIdx.x == 0) {
    incremental_update<RIGID, TRACK>(ctl, p, up, fl, app_step);
    return;
  }
  const int vb = (int)blockIdx.x - fl.ublk;
  int s = 0;
  while (s + 1 < fl.count && vb >= fl.seg[s + 1].blk0) s++;
  const ForceSeg& g = fl.seg[s];
  // blocks go round-robin to the 8 XCDs: within a segment each XCD takes a
  // contiguous range of blocks, so the state of particles shared by
  // neighbouring interactions is fetched into one L2 (a bijection on the
  // segment's blocks)
  const int nseg = (s + 1 < fl.count ? fl.seg[s + 1].blk0 : fl.nblocks) - g.blk0;
  const int b0 = g.blk0 + fl.ublk;  // launch index of the segment's first block
  const int d = ((int)blockIdx.x - b0) & 7;
  const int q8 = nseg >> 3, r8 = nseg & 7;
  const int lbk = d * q8 + min(d, r8) + ((int)blockIdx.x - b0 - d) / 8;
  const int c = lbk * BS + (int)threadIdx.x;
  if (g.kind == ESP_SEG_ROT_BOND) {
    rot_bond_item(ctl, p, c, g, fl);
  } else if (c < g.n + *g.n_app) {
    dyn_item<RIGID>(p, c, g, fl);
  }
}

__global__ __launch_bounds__(BS, 4) void k_forces(StepCtl ctl, PView p, ForceLaunch fl, UpdParams up)
{
  if (step_skip(ctl)) return;
  if (fl.ublk && blockIdx.x == 0) {
    incremental_update<false, false>(ctl, p, up, fl, nullptr);
    return;
  }
  const int vb = (int)blockIdx.x - fl.ublk;
  int s = 0;
  while (s + 1 < fl.count && vb >= fl.seg[s + 1].blk0) s++;
  const ForceSeg& g = fl.seg[s];
  // blocks go round-robin to the 8 XCDs: within a segment each XCD takes a
  // contiguous range of blocks, so the state of particles shared by
  // neighbouring interactions is fetched into one L2 (a bijection on the
  // segment's blocks)
  const int nseg = (s + 1 < fl.count ? fl.seg[s + 1].blk0 : fl.nblocks) - g.blk0;
  const int b0 = g.blk0 + fl.ublk;  // launch index of the segment's first block
  const int d = ((int)blockIdx.x - b0) & 7;
  const int q8 = nseg >> 3, r8 = nseg & 7;
  const int lbk = d * q8 + min(d, r8) + ((int)blockIdx.x - b0 - d) / 8;
  const int c = lbk * BS + (int)threadIdx.x;
  if (g.kind == ESP_SEG_ROT_BOND) {
    rot_bond_item(ctl, p, c, g, fl);
  } else if (c < g.n + *g.n_app) {
    dyn_item<false>(p, c, g, fl);
  }
}

// the same launch with a rigid=True RotFriction group or a tracked group (the
// rigid law needs more registers than the common kernel's budget of 128, so
// this kernel has no minimum occupancy: it runs without register spills)
__global__ __launch_bounds__(BS) void k_forces_ext(StepCtl ctl, PView p, ForceLaunch fl,
                                                      UpdParams up, ForceExt ext)
{
  forces_body<true, true>(ctl, p, fl, up, ext.app_step);
}

__global__ void k_hash_clear(HashRef h)
{
  const unsigned t = blockIdx.x * BS + threadIdx.x;
  if (t <= h.mask) h.t[t] = HS_EMPTY;
}

__global__ void k_hash_insert(HashRef h, const long long* keys, int n)
{
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t < n) hs_insert(h, keys[t]);
}

__global__ void k_bond_keys(int nb, const int2* ij, const int* id, long long* keys)
{
  const int b = blockIdx.x * BS + threadIdx.x;
  if (b >= nb) return;
  const int2 e = ij[b];
  keys[b] = e.x < 0 ? -1ll : pair_key(id[e.x], id[e.y]);
}

__global__ void k_alive_flags(int nb, const int2* ij, int* keep)
{
  const int b = blockIdx.x * BS + threadIdx.x;
  if (b < nb) keep[b] = ij[b].x >= 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// per-particle CSR of force records
// ---------------------------------------------------------------------------
__global__ void k_inc_count(int nc, const int2* ij, int* cnt)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc) return;
  const int2 e = ij[c];
  if (e.x < 0) return;
  atomicAdd(&cnt[e.x], 1);
  atomicAdd(&cnt[e.y], 1);
}

// CSR entry of an interaction for each of its particles: the record reference
// (rec << 2 | has-moments << 1 | second-particle) in the low 32 bits, above it
// the summation order: group position (5 bits) and the interaction's rank in
// the group's reference order (27 bits; the storage index where storage order
// is that order).
__global__ void k_inc_fill(int nc, const int2* ij, int rec_off, int has_m, const int* rank,
                           unsigned long long grp, const int* off, int* fill,
                           unsigned long long* ent)
{
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= nc) return;
  const int2 e = ij[c];
  if (e.x < 0) return;
  const unsigned v = ((unsigned)(rec_off + c) << 2) | ((unsigned)has_m << 1);
  const unsigned long long hi = (grp << 59) | ((unsigned long long)(rank ? rank[c] : c) << 32);
  ent[off[e.x] + atomicAdd(&fill[e.x], 1)] = hi | v;
  ent[off[e.y] + atomicAdd(&fill[e.y], 1)] = hi | v | 1u;
}

// Sorts each particle's run of CSR entries into summation order and keeps the
// record references.  Runs of consecutive particles are contiguous, so a block
// stages its 256 runs in LDS when they fit and each thread insertion-sorts its
// own run there.
#define INC_LDS 6144
__global__ __launch_bounds__(BS) void k_inc_sort(int n, const int* off, unsigned long long* ent64,
                                                 int* ent)
{
  __shared__ unsigned long long buf[INC_LDS];
  const int p0 = blockIdx.x * BS;
  const int p1 = min(p0 + BS, n);
  const int base = off[p0], end = off[p1];
  const int i = p0 + threadIdx.x;
  if (end - base <= INC_LDS) {
    for (int t = base + threadIdx.x; t < end; t += BS) buf[t - base] = ent64[t];
    __syncthreads();
    if (i < n) {
      const int a = off[i] - base, b = off[i + 1] - base;
      for (int s = a + 1; s < b; s++) {
        const unsigned long long v = buf[s];
        int j = s - 1;
        while (j >= a && buf[j] > v) {
          buf[j + 1] = buf[j];
          j--;
        }
        buf[j + 1] = v;
      }
    }
    __syncthreads();
    for (int t = base + threadIdx.x; t < end; t += BS) ent[t] = (int)(unsigned)buf[t - base];
    return;
  }
  if (i >= n) return;
  const int a = off[i], b = off[i + 1];
  for (int s = a + 1; s < b; s++) {
    const unsigned long long v = ent64[s];
    int j = s - 1;
    while (j >= a && ent64[j] > v) {
      ent64[j + 1] = ent64[j];
      j--;
    }
    ent64[j + 1] = v;
  }
  for (int s = a; s < b; s++) ent[s] = (int)(unsigned)ent64[s];
}

__global__ void k_slot_of(int total, const int* ent, int2* slot_of)
{
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= total) return;
  const int v = ent[t];
  int* so = reinterpret_cast<int*>(slot_of + (v >> 2));
  so[v & 1] = t;
}

// ---------------------------------------------------------------------------
// exclusive scan (1024 elements per block, recursive over block sums)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_scan_local(const int* in, int nvalid, int* out, int m, int* bsum)
{
  __shared__ int s[BS];
  const int base = blockIdx.x * 4 * BS + threadIdx.x * 4;
  int v[4];
  int acc = 0;
  for (int k = 0; k < 4; k++) {
    const int idx = base + k;
    v[k] = (idx < nvalid) ? in[idx] : 0;
    acc += v[k];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 1; o < BS; o <<= 1) {
    const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int run = s[threadIdx.x] - acc;  // exclusive prefix of this thread
  for (int k = 0; k < 4; k++) {
    const int idx = base + k;
    if (idx < m) out[idx] = run;
    run += v[k];
  }
  if (bsum && threadIdx.x == BS - 1) bsum[blockIdx.x] = s[BS - 1];
}

__global__ void k_scan_add(int* out, int m, const int* boff)
{
  const int idx = blockIdx.x * 4 * BS + threadIdx.x * 4;
  const int add = boff[blockIdx.x];
  for (int k = 0; k < 4; k++)
    if (idx + k < m) out[idx + k] += add;
}

static size_t scan_tmp_rec(long long m)
{
  const long long nb = (m + 4 * BS - 1) / (4 * BS);
  if (nb <= 1) return 0;
  return (size_t)(2 * nb) + scan_tmp_rec(nb);
}

size_t scan_tmp_size(int n) { return scan_tmp_rec((long long)n + 1) + 16; }

static void scan_rec(hipStream_t s, const int* in, int nvalid, int* out, int m, int* tmp)
{
  const int nb = (m + 4 * BS - 1) / (4 * BS);
  int* bsum = nb > 1 ? tmp : nullptr;
  hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(BS), 0, s, in, nvalid, out, m, bsum);
  if (nb > 1) {
    int* boff = tmp + nb;
    scan_rec(s, bsum, nb, boff, nb, tmp + 2 * nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(BS), 0, s, out, m, boff);
  }
}

void launch_exclusive_scan(hipStream_t s, const int* in, int* out, int n, int* tmp)
{
  scan_rec(s, in, n, out, n + 1, tmp);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int gather_blocks(int n) { return n > 0 ? nblk(n) : 0; }

int force_blocks(int n) { return nblk(n); }

// With a pair of events (ev_a, ev_b both set) the launch stamps them itself at
// the kernel's start and end (hipExtLaunchKernelGGL), so their elapsed time is
// the kernel's own execution time; false when nothing was launched.
bool launch_forces(hipStream_t s, StepCtl ctl, PView p, const ForceLaunch& fl,
                   const UpdParams* up, hipEvent_t ev_a, hipEvent_t ev_b, const ForceExt* ext)
{
  UpdParams u;
  if (up) {
    u = *up;
  } else {
    memset(&u, 0, sizeof u);
  }
  ForceLaunch f = fl;
  f.ublk = up ? 1 : 0;
  if (f.nblocks + f.ublk <= 0) return false;
  bool extended = ext && ext->app_step;
  for (int k = 0; k < f.count; k++)
    extended |= f.seg[k].kind == ESP_SEG_ROT_FRICTION && f.seg[k].rigid;
  ForceExt x;
  memset(&x, 0, sizeof x);
  if (ext) x = *ext;
  if (ev_a && ev_b) {
    if (extended)
      hipExtLaunchKernelGGL(k_forces_ext, dim3(f.nblocks + f.ublk), dim3(BS), 0, s, ev_a, ev_b, 0,
                            ctl, p, f, u, x);
    else
      hipExtLaunchKernelGGL(k_forces, dim3(f.nblocks + f.ublk), dim3(BS), 0, s, ev_a, ev_b, 0, ctl,
                            p, f, u);
  } else if (extended) {
    hipLaunchKernelGGL(k_forces_ext, dim3(f.nblocks + f.ublk), dim3(BS), 0, s, ctl, p, f, u, x);
  } else {
    hipLaunchKernelGGL(k_forces, dim3(f.nblocks + f.ublk), dim3(BS), 0, s, ctl, p, f, u);
  }
  return true;
}

void launch_break_apply(hipStream_t s, StepCtl ctl, PView p, BondSet bs, const int* broken_list,
                        UpdState* st, int* flist, int4* evt, long long step,
                        const GridParams* erase_grid)
{
  if (bs.count <= 0) return;
  if (erase_grid)
    hipLaunchKernelGGL(k_break_apply_erase, dim3(16), dim3(BS), 0, s, ctl, p, bs, broken_list, st,
                       flist, evt, step, *erase_grid);
  else
    hipLaunchKernelGGL(k_break_apply, dim3(16), dim3(BS), 0, s, ctl, p, bs, broken_list, st, flist,
                       evt, step);
}

void launch_hash_clear(hipStream_t s, HashRef h)
{
  hipLaunchKernelGGL(k_hash_clear, dim3(nblk((long long)h.mask + 1)), dim3(BS), 0, s, h);
}

void launch_hash_insert(hipStream_t s, HashRef h, const long long* keys, int n)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hash_insert, dim3(nblk(n)), dim3(BS), 0, s, h, keys, n);
}

void launch_bond_keys(hipStream_t s, int nb, const int2* ij, const int* id, long long* keys)
{
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_bond_keys, dim3(nblk(nb)), dim3(BS), 0, s, nb, ij, id, keys);
}

void launch_alive_flags(hipStream_t s, int nb, const int2* ij, int* keep)
{
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_alive_flags, dim3(nblk(nb)), dim3(BS), 0, s, nb, ij, keep);
}

template <bool STRESS, bool EXTRA>
static void gather_launch(hipStream_t s, StepCtl ctl, PView p, const GatherParams& gp,
                          hipEvent_t ev_a, hipEvent_t ev_b)
{
  const dim3 g(nblk(p.n)), b(BS);
  if (ev_a && ev_b)
    hipExtLaunchKernelGGL((k_gather_integrate<STRESS, EXTRA>), g, b, 0, s, ev_a, ev_b, 0, ctl, p, gp);
  else
    hipLaunchKernelGGL((k_gather_integrate<STRESS, EXTRA>), g, b, 0, s, ctl, p, gp);
}

bool launch_gather_integrate(hipStream_t s, StepCtl ctl, PView p, const GatherParams& gp,
                             hipEvent_t ev_a, hipEvent_t ev_b)
{
  if (p.n <= 0) return false;
  bool extra = gp.ndamp > 0;
  for (int w = 0; w < gp.nwg; w++) extra |= gp.wall[w].bonded != 0 || gp.wall[w].tagged != 0;
  if (gp.sig) {
    if (extra) gather_launch<true, true>(s, ctl, p, gp, ev_a, ev_b);
    else gather_launch<true, false>(s, ctl, p, gp, ev_a, ev_b);
  } else {
    if (extra) gather_launch<false, true>(s, ctl, p, gp, ev_a, ev_b);
    else gather_launch<false, false>(s, ctl, p, gp, ev_a, ev_b);
  }
  return true;
}

// BrittleBeamSCInteraction::broken for every live bond of the stress-criterion
// groups, on the positions after this step's integration and the stresses of
// its forces - the state the reference's next updateInteractions sees.
__global__ __launch_bounds__(BS) void k_sc_break(StepCtl ctl, PView p, ScLaunch sl)
{
  if (step_skip(ctl)) return;
  int g = 0;
  while (g + 1 < sl.count && (int)blockIdx.x >= sl.blk0[g + 1]) g++;
  const int b = ((int)blockIdx.x - sl.blk0[g]) * BS + (int)threadIdx.x;
  if (b >= sl.n[g]) return;
  const int2 e = sl.ij[g][b];
  if (e.x < 0) return;
  const double4 P1 = p.pos[e.x], P2 = x_image(P1, p.pos[e.y], sl.period);
  if (esp_sc_broken(xyz(P1), P1.w, sl.sig + 9 * (size_t)e.x, xyz(P2), P2.w,
                    sl.sig + 9 * (size_t)e.y, sl.sigma3[g], sl.cohesion[g], sl.tan_angle[g])) {
    sl.D4[g][b].w = sl.step;
    const int slot = atomicAdd(&ctl.cur->nbroken, 1);
    sl.broken_list[slot] = sl.bond_off[g] + b;
  }
}

void launch_sc_break(hipStream_t s, StepCtl ctl, PView p, const ScLaunch& sl)
{
  if (sl.nblocks <= 0) return;
  hipLaunchKernelGGL(k_sc_break, dim3(sl.nblocks), dim3(BS), 0, s, ctl, p, sl);
}

void launch_wall_finalize(hipStream_t s, const double* part, int nblocks, int nwg, int nsteps,
                          double* out)
{
  if (nwg <= 0 || nsteps <= 0) return;
  hipLaunchKernelGGL(k_wall_finalize, dim3(nsteps), dim3(BS), 0, s, part, nblocks, nwg, out);
}


void launch_clear_status(hipStream_t s, StepStatus* ring, int ring_size, long long first_step,
                         int count, int* halted)
{
  hipLaunchKernelGGL(k_clear_status, dim3(nblk(count > 0 ? count : 1)), dim3(BS), 0, s, ring,
                     ring_size, first_step, count, halted);
}

void launch_cells_count(hipStream_t s, PView p, GridParams g, int* cell, int* cell_count,
                        int* n_out)
{
  if (p.n <= 0) return;
  hipLaunchKernelGGL(k_cells_count, dim3(nblk(p.n)), dim3(BS), 0, s, p, g, cell, cell_count, n_out);
}

void launch_cells_fill(hipStream_t s, int n, const int* cell, const int* cell_start,
                       int* cell_fill, int* cell_items)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cells_fill, dim3(nblk(n)), dim3(BS), 0, s, n, cell, cell_start, cell_fill,
                     cell_items);
}

void launch_cells_sort(hipStream_t s, int ncell, const int* cell_start, int* cell_items)
{
  if (ncell <= 0) return;
  hipLaunchKernelGGL(k_cells_sort, dim3(nblk(ncell)), dim3(BS), 0, s, ncell, cell_start, cell_items);
}

void launch_pairs(hipStream_t s, int mode_emit, int flagged, PView p, GridParams g, double alpha,
                  const int* cell, const int* cell_start, const int* cell_items, int* cnt,
                  const int* off, int* out_p, int* out_q, long long* out_key)
{
  if (p.n <= 0) return;
  hipLaunchKernelGGL(k_pairs, dim3(nblk(p.n)), dim3(BS), 0, s, mode_emit, flagged, p, g, alpha,
                     cell, cell_start, cell_items, cnt, off, out_p, out_q, out_key);
}

void launch_keep_cells(hipStream_t s, int n, const int* cell, int* keep)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_keep_cells, dim3(nblk(n)), dim3(BS), 0, s, n, cell, keep);
}

void launch_new_slot(hipStream_t s, int n, const int* keep, const int* off, int* nslot)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_new_slot, dim3(nblk(n)), dim3(BS), 0, s, n, keep, off, nslot);
}

template <class T>
static void compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const T* src, T* dst)
{
  if (n <= 0) return;
  hipLaunchKernelGGL((k_compact_field<T>), dim3(nblk(n)), dim3(BS), 0, s, n, width, keep, off, src,
                     dst);
}
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const double4* src, double4* dst)
{
  compact_field(s, n, width, keep, off, src, dst);
}
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const double* src, double* dst)
{
  compact_field(s, n, width, keep, off, src, dst);
}
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const int* src, int* dst)
{
  compact_field(s, n, width, keep, off, src, dst);
}
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const unsigned* src, unsigned* dst)
{
  compact_field(s, n, width, keep, off, src, dst);
}

void launch_remap_pairs(hipStream_t s, int nc, const int2* ij, const int* nslot, int2* out,
                        int* keep)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_remap_pairs, dim3(nblk(nc)), dim3(BS), 0, s, nc, ij, nslot, out, keep);
}

void launch_pack_dyn(hipStream_t s, int nc, const int* keep, const int* off, const int2* ij,
                     const double4* ff, int2* oij, long long* okey, double4* off_)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_pack_dyn, dim3(nblk(nc)), dim3(BS), 0, s, nc, keep, off, ij, ff, oij, okey,
                     off_);
}

void launch_wall_cands(hipStream_t s, PView p, int w, const double4* init, int cap, int* cnt,
                       int* slot, unsigned* flg, double4* pos, double4* ipos)
{
  if (p.n <= 0) return;
  hipLaunchKernelGGL(k_wall_cands, dim3(nblk(p.n)), dim3(BS), 0, s, p, w, init, cap, cnt, slot,
                     flg, pos, ipos);
}

void launch_wall_mark_ordered(hipStream_t s, int no, const int* ord_slot, const unsigned* flags,
                              int w, int* mark)
{
  if (no <= 0) return;
  hipLaunchKernelGGL(k_wall_mark_ordered, dim3(nblk(no)), dim3(BS), 0, s, no, ord_slot, flags, w,
                     mark);
}

void launch_wall_pack_ordered(hipStream_t s, int no, const int* ord_slot, const int* mark,
                              const int* off, PView p, const double4* init, WallCand* out)
{
  if (no <= 0) return;
  hipLaunchKernelGGL(k_wall_pack_ordered, dim3(nblk(no)), dim3(BS), 0, s, no, ord_slot, mark, off,
                     p, init, out);
}

void launch_fill_int(hipStream_t s, int n, int* a, int v)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(n)), dim3(BS), 0, s, n, a, v);
}

void launch_clear_flag_bit(hipStream_t s, PView p, UpdState* st)
{
  hipLaunchKernelGGL(k_clear_flag_bit, dim3(nblk(p.n > 0 ? p.n : 1)), dim3(BS), 0, s, p, st);
}

void launch_ed_mark(hipStream_t s, int nc, const int* cp, const int* cq, const long long* ckey,
                    ExclSet ex, const long long* okey, int nold, const int2* oij, PView p,
                    double period, int* keep, int* src)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_ed_mark, dim3(nblk(nc)), dim3(BS), 0, s, nc, cp, cq, ckey, ex, okey, nold,
                     oij, p, period, keep, src);
}

__global__ void k_wrap_x(PView p, double x0, double period, double4* init)
{
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= p.n) return;
  double4 P = p.pos[i];
  double sh = 0.0;
  if (P.x < x0) sh = period;
  else if (P.x >= x0 + period) sh = -period;
  if (sh == 0.0) return;
  P.x = P.x + sh;
  p.pos[i] = P;
  double4 O = p.old[i];
  O.x = O.x + sh;
  p.old[i] = O;
  if (init) {
    double4 I = init[i];
    I.x = I.x + sh;
    init[i] = I;
  }
}

void launch_wrap_x(hipStream_t s, PView p, double x0, double period, double4* init)
{
  if (p.n <= 0) return;
  hipLaunchKernelGGL(k_wrap_x, dim3(nblk(p.n)), dim3(BS), 0, s, p, x0, period, init);
}

void launch_ne_mark(hipStream_t s, int nc, const int* cp, const int* cq, const int* id, ExclSet ex,
                    int* keep)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_ne_mark, dim3(nblk(nc)), dim3(BS), 0, s, nc, cp, cq, id, ex, keep);
}

void launch_pair_emit(hipStream_t s, int nc, const int* keep, const int* off, const int* cp,
                      const int* cq, const long long* ckey, const int* src, const double4* off_ff,
                      int2* nij, long long* nkey, double4* nff)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_pair_emit, dim3(nblk(nc)), dim3(BS), 0, s, nc, keep, off, cp, cq, ckey, src,
                     off_ff, nij, nkey, nff);
}

void launch_filter_new(hipStream_t s, int nc, const int* cp, const int* cq, const int* id,
                       const long long* ckey, ExclSet ex, HashRef present, int check_present,
                       int* keep)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_filter_new, dim3(nblk(nc)), dim3(BS), 0, s, nc, cp, cq, id, ckey, ex,
                     present, check_present, keep);
}

void launch_merge(hipStream_t s, int nold, const int2* oij, const long long* okey,
                  const double4* off_, int nnew, const int2* nij_in, const long long* nkey_in,
                  const double4* nff_in, int2* ij_out, long long* key_out, double4* ff_out)
{
  if (nold + nnew <= 0) return;
  hipLaunchKernelGGL(k_merge, dim3(nblk((long long)nold + nnew)), dim3(BS), 0, s, nold, oij, okey,
                     off_, nnew, nij_in, nkey_in, nff_in, ij_out, key_out, ff_out);
}

void launch_pack_pairs(hipStream_t s, int n, const int* keep, const int* off, const int* cp,
                       const int* cq, const long long* ckey, int2* ij, long long* key)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pack_pairs, dim3(nblk(n)), dim3(BS), 0, s, n, keep, off, cp, cq, ckey, ij, key);
}


void launch_compact_bonds(hipStream_t s, int nb, const int* keep, const int* off,
                          const int2* ij, const double4* k4, const double4* mx4,
                          const double4* D4, const int* rank, int2* ij2, double4* k42,
                          double4* mx42, double4* D42, int* rank2)
{
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_compact_bonds, dim3(nblk(nb)), dim3(BS), 0, s, nb, keep, off, ij, k4, mx4,
                     D4, rank, ij2, k42, mx42, D42, rank2);
}



void launch_wall_update(hipStream_t s, PView p, GridParams g, int nwg, const WallParam* w)
{
  if (p.n <= 0) return;
  WallSet ws;
  for (int k = 0; k < ESP_MAX_WALL_GROUPS; k++) ws.w[k] = w[k];
  hipLaunchKernelGGL(k_wall_update, dim3(nblk(p.n)), dim3(BS), 0, s, p, g, nwg, ws);
}

void launch_inc_count(hipStream_t s, int nc, const int2* ij, int* inc_cnt)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_inc_count, dim3(nblk(nc)), dim3(BS), 0, s, nc, ij, inc_cnt);
}

void launch_inc_fill(hipStream_t s, int nc, const int2* ij, int rec_off, int has_m, const int* rank,
                     int grp, const int* inc_off, int* inc_fill, unsigned long long* ent64)
{
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_inc_fill, dim3(nblk(nc)), dim3(BS), 0, s, nc, ij, rec_off, has_m, rank,
                     (unsigned long long)grp, inc_off, inc_fill, ent64);
}

void launch_slot_of(hipStream_t s, int total, const int* inc_ent, int2* slot_of)
{
  if (total <= 0) return;
  hipLaunchKernelGGL(k_slot_of, dim3(nblk(total)), dim3(BS), 0, s, total, inc_ent, slot_of);
}

void launch_inc_sort(hipStream_t s, int n, const int* inc_off, unsigned long long* ent64,
                     int* inc_ent)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_inc_sort, dim3(nblk(n)), dim3(BS), 0, s, n, inc_off, ent64, inc_ent);
}
