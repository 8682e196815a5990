/*
 * esp_oracle.c - CPU restatement of ESyS-Particle's per-timestep particle engine.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): the checker the HIP engine is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; the product library never links or calls it.
 *
 * What it restates (paths relative to the ESyS-Particle 2.3.5 tree), for one
 * worker (1x1x1 process grid, so every MPI exchange is a no-op):
 *   - Vec3 / Matrix3 / Quaternion arithmetic with the reference's expression
 *     trees (Foundation/vec3.hpp:80-244, Matrix3.hpp:31-140, Quaternion.hpp:162-309);
 *     the build uses -O2 -ffp-contract=off so no FMA is formed, as in the
 *     reference's x86-64 -O2 build.
 *   - NeighborTable index/insert/build/getFullList/getNewList/getParticlesAtPlane/
 *     isInInner (ntable/src/ntable.hpp:128-288, 471-690) including the
 *     std::list / cell-vector orders, which fix the order in which interactions
 *     are stored and forces are summed.
 *   - ParallelParticleArray construction and rebuild timestamp (ppa/src/pp_array.hpp:61-103,
 *     433-554) and TSubLattice::initNeighborTable (Parallel/SubLattice.hpp:455-478).
 *   - Interaction storages: PIS_E rebuild/tryInsert (pis/pi_storage_e.hpp:111-250),
 *     PIS_NE update (pis/pi_storage_ne.hpp:49-102), PIS_ED update/addExIG
 *     (pis/pi_storage_ed.hpp:20-161), PIS_EB update (pis/pi_storage_eb.hpp:23-49).
 *   - Laws: CElasticInteraction (Model/ElasticInteraction.cpp:25-71),
 *     CRotElasticInteraction (Model/RotElasticInteraction.cpp:40-88),
 *     CRotFrictionInteraction (Model/RotFricInteraction.cpp:103-231, 345-351),
 *     ARotBondedInteraction (Model/ARotBondedInteraction.cpp:22-51, 141-180, 307-448),
 *     CRotBondedInteraction (Model/RotBondedInteraction.cpp:122-219),
 *     CElasticWallInteraction / CEWallInteractionGroup (Model/EWallInteraction.hpp:26-78,
 *     Model/EWallInteractionGroup.hpp:42-82), BodyForceGroup (Model/BodyForceGroup.hpp:41-65).
 *   - Particles: CParticle (Model/Particle.cpp:57-122, 257-283, 391-400),
 *     CRotParticle (Model/RotParticle.cpp:80-177, 277-332).
 *   - Step order: CLatticeMaster::runOneStep (Parallel/LatticeMaster.cpp:2602-2725)
 *     with TSubLattice::checkNeighbors / searchNeighbors / rebuildInteractions /
 *     updateInteractions / zeroForces / calcForces / integrate
 *     (Parallel/SubLattice.hpp:2066-2236, 2354-2536).
 *
 * Parity of this restatement is pinned by tests/test_oracle_*.py (closed forms,
 * brute-force neighbour sets, invariants and the reference's own Foundation
 * math compiled from source into oracle/_ref).
 */
#define _GNU_SOURCE
#include "esp_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Foundation math (vec3.hpp, Matrix3.hpp, Quaternion.hpp)                     */
/* ------------------------------------------------------------------------- */
typedef struct { double x, y, z; } v3;
typedef struct { double s; v3 v; } quat;
typedef struct { double m[3][3]; } mat3;

static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
/* Vec3*double and double*Vec3 (vec3.hpp:118-121, 221-224): both round s*x. */
static inline v3 vmul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
/* dot / operator* (vec3.hpp:113-116, 212-217): left-associated sum. */
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vnorm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline double vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* Matrix3::operator*(Vec3) (Matrix3.hpp:126-133). */
static inline v3 mmul(const mat3* m, v3 v)
{
  return V(m->m[0][0] * v.x + m->m[0][1] * v.y + m->m[0][2] * v.z,
           m->m[1][0] * v.x + m->m[1][1] * v.y + m->m[1][2] * v.z,
           m->m[2][0] * v.x + m->m[2][1] * v.y + m->m[2][2] * v.z);
}
static inline mat3 mtrans(const mat3* a)
{
  mat3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = a->m[j][i];
  return r;
}
/* Quaternion::operator*(Quaternion) (Quaternion.hpp:178-195). */
static inline quat qmul(quat a, quat b)
{
  quat r;
  r.s = a.s * b.s - vdot(a.v, b.v);
  r.v = vadd(vadd(vmul(b.v, a.s), vmul(a.v, b.s)), vcross(a.v, b.v));
  return r;
}
static inline quat qinv(quat a) { quat r = {a.s, vneg(a.v)}; return r; }
/* Quaternion::to_matrix (Quaternion.hpp:273-309). */
static inline mat3 qmat(quat q)
{
  mat3 r;
  double s = q.s, x = q.v.x, y = q.v.y, z = q.v.z;
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

/* ------------------------------------------------------------------------- */
/* Growable arrays                                                             */
/* ------------------------------------------------------------------------- */
#define GROW(arr, n, cap)                                                     \
  do {                                                                        \
    if ((n) >= (cap)) {                                                       \
      size_t nc_ = (cap) ? 2 * (size_t)(cap) : 16;                            \
      while (nc_ <= (size_t)(n)) nc_ *= 2;                                    \
      void* p_ = realloc((arr), nc_ * sizeof(*(arr)));                        \
      if (!p_) abort();                                                       \
      (arr) = p_;                                                             \
      (cap) = (int)nc_;                                                       \
    }                                                                         \
  } while (0)

/* ------------------------------------------------------------------------- */
/* Engine state                                                                */
/* ------------------------------------------------------------------------- */
typedef struct {
  int id, tag;
  v3 pos, vel, force, oldpos, initpos, angvel, moment;
  double rad, mass, inv_mass, inert, inv_inert;
  quat q;
  double sigma[3][3];
  /* sums of the norms of the forces / moments applied in the last step: the
   * per-particle scale of the parity tolerance (test instrumentation) */
  double fabs_sum, mabs_sum;
  int flag, is_dyn, is_rot;
} opart;

enum { G_NE_ELASTIC = 1, G_NE_ROTELASTIC = 2, G_ED_ROTFRICTION = 3 };

typedef struct {
  int id1, id2, i1, i2;
  double k;               /* scaled normal stiffness (elastic / friction / rot elastic) */
  double ks;              /* scaled shear stiffness (friction) */
  v3 ffric;               /* shear-spring force (friction) */
  int slipping, touching; /* friction state */
  double ediss;
} odyn;

typedef struct {
  char name[128];
  int kind;
  double k, ks, mu_d, mu_s; /* group parameters (unscaled) */
  int scaling, meanR_scaling;
  odyn* v;
  int n, cap;
  long long update_ts;
  int ex[16]; /* indices of excluding bonded groups */
  int nex;
  /* createInteractionGroupTagged: tags sorted so that tag1 <= tag2
   * (pi_storage_ed_t.hpp:15-30, pi_storage_ne_t.hpp:15-30) */
  int tagged, tag1, mask1, tag2, mask2;
  int rigid; /* RotFriction rigid=True: calcRigidBodyForces */
  /* pair-event log (eso_track_pair_events): the stored pairs as sorted id
   * keys after the last step, and the events, 4 ints each (step, id1, id2, kind) */
  int track;
  long long* evprev;
  int nevprev;
  int* ev;
  int nev, capev;
} ogroup;

typedef struct {
  int id1, id2, i1, i2;
  double kr, ks, kt, kb, max_n, max_s, max_t, max_b;
  v3 D;
  double dist;
  double nForce, shForce, tMoment, bMoment;
} obond;

typedef struct {
  char name[128];
  esp_rot_bond_prms prm;
  /* BrittleBeamSC groups: stress criterion instead of CRotBondedInteraction::broken */
  int sc;
  double sc_sigma3, sc_cohesion, sc_tan;
  obond* v;
  int n, cap;
  long long* keys; /* sorted (id1,id2) keys of live bonds: the storage's m_set */
  int nkeys, capkeys;
  int keys_dirty;
} obgroup;

typedef struct {
  char name[128];
  v3 origin, normal, force, step_motion;
  v3 oldpos; /* CWall::m_oldpos: the origin at creation (Wall.cpp:38) */
  int has_motion;
  double* hist; /* force after each step */
  int nhist, caphist;
} owall;

typedef struct {
  char name[128];
  int wall;
  double k;
  int bonded, tag, mask; /* CBWallInteractionGroup (BWallInteractionGroup.hpp) */
  int tagged_only;       /* CTaggedEWallInteractionGroup (TaggedEWallInteractionGroup.hpp) */
  int* pidx; /* candidate particles (list order) */
  int* inner;
  double* keff;
  int* tagged; /* bonded groups: the candidate's tag matches (bonded, not elastic) */
  int n, cap;
} owgroup;

typedef struct {
  char name[128];
  v3 g;
} ograv;

/* a damping group (TSubLattice::doAddDamping, SubLattice.hpp:1747-1783,
 * RotSubLattice.hpp:78-112): key = its key in the damping map, "Damping" /
 * "RotDamping" for those types, the group name for the local forms */
typedef struct {
  char key[128];
  int type;
  double visc;
  int maxit;
} odamp;

struct eso_engine {
  esp_config cfg;
  int rot, do2d;
  double dt;
  /* neighbour table (ntable.hpp:52-85) */
  v3 p0;
  double dim, alpha;
  /* circular x (TSubLattice::initNeighborTable with circ, SubLattice.hpp:495-564):
   * period L = table length in x, cells 0..nx-1 without padding in x */
  int circ_x;
  double period;
  int gidx, gidy, gidz, nx, ny, nz, ncell;
  int* cell_start; /* CSR of cell contents in list order */
  int* cell_items;
  /* particles in std::list order */
  opart* P;
  int np, capp;
  long long* idkey; /* sorted (id<<32 | index) for id lookup */
  int nidkey, capidkey;
  /* connections per tag, in arrival order */
  int* conn; /* triples tag,id1,id2 */
  int nconn, capconn;
  /* groups */
  ogroup* dg;
  int ndg, capdg;
  obgroup* bg;
  int nbg, capbg;
  owall* wl;
  int nwl, capwl;
  owgroup* wg;
  int nwg, capwg;
  ograv* gv;
  int ngv, capgv;
  odamp dp[4];
  int ndp;
  /* step bookkeeping */
  long long t, timestamp, last_ns;
  int first_time;
  long long num_host_phases;
  /* logs */
  int* brk; /* triples step,id1,id2 */
  int nbrk, capbrk;
  int* rebuild_steps;
  int nreb, capreb;
  long long* last_full; /* sorted keys of the last full list */
  int nlast, caplast;
  char err[512];
};

static char g_create_err[512];

static int fail(eso_engine* e, int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(e->err, sizeof e->err, fmt, ap);
  va_end(ap);
  return code;
}

static inline long long pkey(int a, int b)
{
  return ((long long)(unsigned)a << 32) | (long long)(unsigned)b;
}

static int cmp_ll(const void* a, const void* b)
{
  long long x = *(const long long*)a, y = *(const long long*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

typedef struct { long long k; int i; } keyidx;

/* stable bottom-up merge sort of (key, index) pairs */
static void sort_keyidx(keyidx* a, int n)
{
  keyidx* tmp = malloc(sizeof(keyidx) * (size_t)(n > 0 ? n : 1));
  keyidx* src = a;
  keyidx* dst = tmp;
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int x = lo, y = mid, o = lo;
      while (x < mid && y < hi) dst[o++] = (src[y].k < src[x].k) ? src[y++] : src[x++];
      while (x < mid) dst[o++] = src[x++];
      while (y < hi) dst[o++] = src[y++];
    }
    keyidx* t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(keyidx) * (size_t)n);
  free(tmp);
}

static int bsearch_ll(const long long* arr, int n, long long key)
{
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (arr[mid] == key) return mid;
    if (arr[mid] < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* ------------------------------------------------------------------------- */
/* Neighbour table                                                             */
/* ------------------------------------------------------------------------- */
/* floor((x - x0)/dim) as the reference computes it, with x86's conversion of
 * out-of-range / NaN values (INT_MIN) mapped to "outside". */
static inline int cell_coord(double x, double x0, double dim, int* ok)
{
  double f = floor((x - x0) / dim);
  if (!(f >= -2147483648.0 && f < 2147483648.0)) { *ok = 0; return 0; }
  return (int)f;
}

/* NeighborTable::index(const Vec3&) (ntable.hpp:128-139). */
static int nt_index(const eso_engine* e, v3 pos)
{
  int ok = 1;
  int ix = cell_coord(pos.x, e->p0.x, e->dim, &ok) - e->gidx;
  /* circular x: positions are wrapped into [x0, x0 + L) at each rebuild; a
   * rounding of (x - x0)/dim onto the end of the range stays in the last cell */
  if (e->circ_x && ok && pos.x >= e->p0.x && pos.x < e->p0.x + e->period)
    ix = ix < 0 ? 0 : (ix >= e->nx ? e->nx - 1 : ix);
  int iy = cell_coord(pos.y, e->p0.y, e->dim, &ok) - e->gidy;
  int iz = cell_coord(pos.z, e->p0.z, e->dim, &ok) - e->gidz;
  if (!ok) return -1;
  if (ix >= 0 && ix < e->nx && iy >= 0 && iy < e->ny && iz >= 0 && iz < e->nz)
    return e->ny * e->nz * ix + e->nz * iy + iz;
  return -1;
}

/* NeighborTable::isInInner (ntable.hpp:207-222).  With circular x the table
 * has no padding cells in x, so every x is inner. */
static int nt_inner(const eso_engine* e, v3 pos)
{
  int ok = 1;
  int ix = cell_coord(pos.x, e->p0.x, e->dim, &ok) - e->gidx;
  int iy = cell_coord(pos.y, e->p0.y, e->dim, &ok) - e->gidy;
  int iz = cell_coord(pos.z, e->p0.z, e->dim, &ok) - e->gidz;
  if (!ok) return 0;
  int xin = e->circ_x ? 1 : (ix > 0) && (ix < e->nx - 1);
  if (e->nz < 3) return xin && (iy > 0) && (iy < e->ny - 1);
  return xin && (iy > 0) && (iy < e->ny - 1) && (iz > 0) && (iz < e->nz - 1);
}

/* The image of x2 nearest to x1 in a circular x dimension: x2 shifted by the
 * period L when the two are more than L/2 apart, one addition, as a ghost copy
 * carries the shifted position (CParticle::setCircular, Particle.cpp:336-342).
 * The particle of higher id is the one imaged to the other's side. */
static inline v3 x_image(const eso_engine* e, v3 x1, v3 x2)
{
  if (e->circ_x) {
    double d = x1.x - x2.x;
    if (d > 0.5 * e->period) x2.x = x2.x + e->period;
    else if (d < -0.5 * e->period) x2.x = x2.x - e->period;
  }
  return x2;
}

static void rebuild_idmap(eso_engine* e)
{
  e->nidkey = 0;
  GROW(e->idkey, e->np, e->capidkey);
  for (int i = 0; i < e->np; i++)
    e->idkey[e->nidkey++] = ((long long)(unsigned)e->P[i].id << 32) | (unsigned)i;
  qsort(e->idkey, e->nidkey, sizeof(long long), cmp_ll);
}

static int find_particle(const eso_engine* e, int id)
{
  int lo = 0, hi = e->nidkey - 1;
  unsigned uid = (unsigned)id;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    unsigned k = (unsigned)(e->idkey[mid] >> 32);
    if (k == uid) return (int)(e->idkey[mid] & 0xffffffffLL);
    if (k < uid) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* Cell contents in list order: NeighborTable::build (ntable.hpp:262-288) after
 * the out-of-grid particles have been removed. */
static void rebuild_cells(eso_engine* e)
{
  free(e->cell_start);
  free(e->cell_items);
  e->cell_start = calloc((size_t)e->ncell + 1, sizeof(int));
  e->cell_items = malloc(sizeof(int) * (size_t)(e->np > 0 ? e->np : 1));
  int* cidx = malloc(sizeof(int) * (size_t)(e->np > 0 ? e->np : 1));
  for (int i = 0; i < e->np; i++) {
    cidx[i] = nt_index(e, e->P[i].pos);
    e->cell_start[cidx[i] + 1]++;
  }
  for (int c = 0; c < e->ncell; c++) e->cell_start[c + 1] += e->cell_start[c];
  int* fill = calloc((size_t)e->ncell, sizeof(int));
  for (int i = 0; i < e->np; i++) {
    int c = cidx[i];
    e->cell_items[e->cell_start[c] + fill[c]++] = i;
  }
  free(fill);
  free(cidx);
}

/* Interactions keep particle ids and are re-pointed by id after the particle
 * list changes (PIS_E::rebuild, pi_storage_e.hpp:111-155); interactions whose
 * particles are gone are erased, keeping the list order of the rest. */
static void repoint_interactions(eso_engine* e)
{
  for (int g = 0; g < e->nbg; g++) {
    obgroup* G = &e->bg[g];
    int w = 0;
    for (int i = 0; i < G->n; i++) {
      obond b = G->v[i];
      b.i1 = find_particle(e, b.id1);
      b.i2 = find_particle(e, b.id2);
      if (b.i1 >= 0 && b.i2 >= 0) G->v[w++] = b;
      else G->keys_dirty = 1;
    }
    G->n = w;
  }
  for (int g = 0; g < e->ndg; g++) {
    ogroup* G = &e->dg[g];
    int w = 0;
    for (int i = 0; i < G->n; i++) {
      odyn d = G->v[i];
      d.i1 = find_particle(e, d.id1);
      d.i2 = find_particle(e, d.id2);
      if (d.i1 >= 0 && d.i2 >= 0) G->v[w++] = d;
    }
    G->n = w;
  }
  for (int g = 0; g < e->nwg; g++) e->wg[g].n = 0; /* refilled by the update */
}

/* NeighborTable::build (ntable.hpp:262-288): particles outside the padded grid
 * are erased from the list. */
static void nt_build(eso_engine* e)
{
  int w = 0, removed = 0;
  /* circular x: a particle that crossed an x end moves to its image on the
   * other side, with its initial and old positions (CParticle::setCircular,
   * Particle.cpp:336-342, applied when the ghost copy takes the particle over
   * at the table rebuild, pp_array.hpp:433-493) */
  if (e->circ_x)
    for (int i = 0; i < e->np; i++) {
      opart* p = &e->P[i];
      double sh = 0.0;
      if (p->pos.x < e->p0.x) sh = e->period;
      else if (p->pos.x >= e->p0.x + e->period) sh = -e->period;
      if (sh != 0.0) {
        p->pos.x = p->pos.x + sh;
        p->initpos.x = p->initpos.x + sh;
        p->oldpos.x = p->oldpos.x + sh;
      }
    }
  for (int i = 0; i < e->np; i++) {
    if (nt_index(e, e->P[i].pos) != -1) e->P[w++] = e->P[i];
    else removed++;
  }
  e->np = w;
  rebuild_idmap(e);
  rebuild_cells(e);
  if (removed) repoint_interactions(e);
}

typedef struct { int* a; int* b; int n, cap; } pairlist; /* particle indices (lower id first) */

static void pl_push(pairlist* L, const eso_engine* e, int i, int j)
{
  if (L->n >= L->cap) {
    L->cap = L->cap ? 2 * L->cap : 1024;
    L->a = realloc(L->a, sizeof(int) * (size_t)L->cap);
    L->b = realloc(L->b, sizeof(int) * (size_t)L->cap);
    if (!L->a || !L->b) abort();
  }
  if (e->P[i].id < e->P[j].id) { L->a[L->n] = i; L->b[L->n] = j; }
  else { L->a[L->n] = j; L->b[L->n] = i; }
  L->n++;
}

static inline int nt_lin(const eso_engine* e, int x, int y, int z)
{
  return e->ny * e->nz * x + e->nz * y + z;
}

/* addPairsToListLocal / addPairsToList (ntable.hpp:471-521). */
static inline int near_pair(const eso_engine* e, int i, int j)
{
  /* from the particle of lower id (the same norm without a circular image) */
  if (e->P[j].id < e->P[i].id) { int t = i; i = j; j = t; }
  double dist2 = vnorm2(vsub(e->P[i].pos, x_image(e, e->P[i].pos, e->P[j].pos)));
  double dmax = e->P[i].rad + e->P[j].rad + e->alpha;
  return dist2 <= (dmax * dmax);
}

/* getFullList (ntable.hpp:590-621) / getNewList (ntable.hpp:628-664). */
static void nt_pairs(eso_engine* e, pairlist* L, int flagged)
{
  L->n = 0;
  for (int ix = 0; ix < e->nx; ix++)
    for (int iy = 0; iy < e->ny; iy++)
      for (int iz = 0; iz < e->nz; iz++) {
        int idx1 = nt_lin(e, ix, iy, iz);
        int s1 = e->cell_start[idx1], n1 = e->cell_start[idx1 + 1] - s1;
        if (n1 >= 2) {
          for (int a = 0; a < n1 - 1; a++) {
            int i = e->cell_items[s1 + a];
            if (flagged && !e->P[i].flag) continue;
            for (int b = a + 1; b < n1; b++) {
              int j = e->cell_items[s1 + b];
              if (flagged ? e->P[j].flag : near_pair(e, i, j)) pl_push(L, e, i, j);
            }
          }
        }
        int xmax = (ix < e->nx - 1) ? ix + 1 : ix;
        int ymax = (iy < e->ny - 1) ? iy + 1 : iy;
        int zmax = (iz < e->nz - 1) ? iz + 1 : iz;
        int xmin = (ix > 0) ? ix - 1 : ix;
        int ymin = (iy > 0) ? iy - 1 : iy;
        int zmin = (iz > 0) ? iz - 1 : iz;
        if (e->circ_x) { /* neighbours across the circular x edge (nx >= 3) */
          xmin = ix - 1;
          xmax = ix + 1;
        }
        for (int x = xmin; x <= xmax; x++)
          for (int y = ymin; y <= ymax; y++)
            for (int z = zmin; z <= zmax; z++) {
              int idx2 = nt_lin(e, e->circ_x ? (x + e->nx) % e->nx : x, y, z);
              if (idx2 <= idx1) continue;
              int s2 = e->cell_start[idx2], n2 = e->cell_start[idx2 + 1] - s2;
              for (int a = 0; a < n1; a++) {
                int i = e->cell_items[s1 + a];
                if (flagged && !e->P[i].flag) continue;
                for (int b = 0; b < n2; b++) {
                  int j = e->cell_items[s2 + b];
                  if (flagged ? e->P[j].flag : near_pair(e, i, j)) pl_push(L, e, i, j);
                }
              }
            }
      }
  if (flagged)
    for (int i = 0; i < e->np; i++) e->P[i].flag = 0;
}

/* ------------------------------------------------------------------------- */
/* Particles                                                                   */
/* ------------------------------------------------------------------------- */
/* CParticle::applyForce (Particle.cpp:391-400). */
static inline void apply_force_at(opart* p, v3 f, v3 pos, v3 ppos)
{
  p->force = vadd(p->force, f);
  p->fabs_sum += vnorm(f);
  v3 rel = vsub(pos, ppos);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) p->sigma[i][j] -= vget(f, i) * vget(rel, j);
}
/* the contact point relative to the particle's own position; an interaction
 * across the circular edge passes the position of the image it computed with */
static inline void apply_force(opart* p, v3 f, v3 pos)
{
  p->force = vadd(p->force, f);
  p->fabs_sum += vnorm(f);
  v3 rel = vsub(pos, p->pos);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) p->sigma[i][j] -= vget(f, i) * vget(rel, j);
}
/* CRotParticle::applyMoment (RotParticle.cpp:329-332). */
static inline void apply_moment(opart* p, v3 m)
{
  p->moment = vadd(p->moment, m);
  p->mabs_sum += vnorm(m);
}

/* CRotParticle::integrate + CParticle::integrate + CRotParticle::rescale
 * (RotParticle.cpp:277-321, Particle.cpp:257-274). */
static void integrate_particle(opart* p, double dt, int rot, int do2d)
{
  if (rot) {
    if (p->is_rot) {
      if (do2d) {
        p->force = V(p->force.x, p->force.y, 0);
        p->moment = V(0, 0, p->moment.z);
      }
      p->angvel = vadd(p->angvel, vmul(p->moment, dt * p->inv_inert));
      quat w = {0, p->angvel};
      quat dq = qmul(w, p->q);
      double h = dt / 2.0;
      p->q.s += h * dq.s;
      p->q.v = vadd(p->q.v, vmul(dq.v, h));
    } else {
      p->angvel = V(0.0, 0.0, 0.0);
    }
  }
  if (p->is_dyn) {
    if (do2d) p->force = V(p->force.x, p->force.y, 0);
    p->vel = vadd(p->vel, vmul(vmul(p->force, p->inv_mass), dt));
    p->pos = vadd(p->pos, vmul(p->vel, dt));
  }
  if (rot) {
    double module = sqrt(p->q.s * p->q.s + p->q.v.x * p->q.v.x + p->q.v.y * p->q.v.y +
                         p->q.v.z * p->q.v.z);
    if (module != 0.0) {
      double inverse = 1.0 / module;
      p->q.s = inverse * p->q.s;
      p->q.v = vmul(p->q.v, inverse);
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Interaction laws                                                            */
/* ------------------------------------------------------------------------- */
/* calc_angle (ARotBondedInteraction.cpp:22-51). */
static double calc_angle(double s_in, double c_os)
{
  double angle;
  if (s_in > 0.0) {
    if (c_os >= 1.0 || c_os <= -1.0) angle = 0.0;
    else angle = acos(c_os);
  } else if (s_in == 0.0) {
    angle = 0.0;
  } else {
    if (c_os >= 1.0 || c_os <= -1.0) angle = 0.0;
    else angle = -acos(c_os);
  }
  return angle;
}

typedef struct { v3 f1, m1, f2, m2, cpos; double nForce, shForce, tMoment, bMoment; } bond_out;

/* ARotBondedInteraction::calcForces (ARotBondedInteraction.cpp:307-448). */
static void rot_bond_law(v3 pos1, quat q1, double r1, v3 pos2, quat q2, double r2,
                         double kr, double ks, double kt, double kb, v3 Dinit, bond_out* o)
{
  const double HALF_SQRT_2 = 0.5 * sqrt(2.0);
  double s_fai = 0.0, c_fai = 0.0, c_pasi2 = 0.0, s_pasi2 = 0.0;
  double sita = 0.0, pasi = 0.0;

  const mat3 mat0 = qmat(q1);
  const v3 rbp = vsub(pos2, pos1);
  const double rbpNorm = vnorm(rbp);
  const v3 rb = mmul(&mat0, vsub(pos2, pos1));
  const double rbNorm = vnorm(rb);
  const double r_0Norm = r1 + r2;
  const v3 delta_r = vdiv(vmul(rb, rbNorm - r_0Norm), rbNorm);
  const v3 Fr = vmul(delta_r, kr);

  const double gama_cos = vdot(rb, Dinit) / (rbNorm * vnorm(Dinit));
  double gama = 0.0;
  if (gama_cos < 1.0 && gama_cos > -1.0) gama = acos(gama_cos);

  const v3 rb_0 = vcross(rb, Dinit);
  const v3 rb_b0 = vcross(rb, rb_0);
  const double rb_b0Norm = vnorm(rb_b0);
  v3 Fst = V(0.0, 0.0, 0.0);
  if (rb_b0Norm != 0) Fst = vdiv(vmul(rb_b0, ks * r_0Norm * gama), rb_b0Norm);

  v3 Mst = V(0.0, 0.0, 0.0);
  const double rb_0Norm = vnorm(rb_0);
  if (rb_0Norm != 0) Mst = vdiv(vmul(rb_0, -vnorm(Fst)), rb_0Norm);

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

  const quat qm = {m0, {m1, m2, m3}};
  const mat3 mat = qmat(qm);
  const mat3 matTrans = mtrans(&mat);
  const quat rp = qmul(qinv(q1), q2);
  const quat r = qmul(qmul(qinv(qm), rp), qm);
  const double temp0 = r.s * r.s + r.v.z * r.v.z;
  if (temp0 == 0.0) {
    pasi = 0.0;
  } else {
    const double sqrtTemp0 = sqrt(temp0);
    c_pasi2 = r.s / sqrtTemp0;
    s_pasi2 = r.v.z / sqrtTemp0;
    pasi = 2.0 * calc_angle(s_pasi2, c_pasi2);
  }
  const v3 Mtp = V(0.0, 0.0, kt * pasi);
  const v3 Mtr = mmul(&matTrans, Mtp);
  const double temp2 = r.v.x * r.v.x + r.v.y * r.v.y;
  const double temp3 = r.v.x * r.v.z + r.s * r.v.y;
  const double temp4 = r.v.y * r.v.z - r.s * r.v.x;

  v3 Mbr = V(0, 0, 0), Fsr = V(0, 0, 0), Msr = V(0, 0, 0);
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
    const v3 Mbp = V(-0.5 * kb * sita * s_fai, 0.5 * kb * sita * c_fai, 0.0);
    Mbr = mmul(&matTrans, Mbp);
    const v3 Fsp = V(-0.5 * ks * c_fai * sita * r_0Norm, -0.5 * ks * s_fai * sita * r_0Norm, 0.0);
    Fsr = mmul(&matTrans, Fsp);
    const v3 Msp = V(0.5 * ks * s_fai * sita * r_0Norm, -0.5 * ks * c_fai * sita * r_0Norm, 0.0);
    Msr = mmul(&matTrans, Msp);
  }

  const double eq_rad1 = r1 / (r1 + r2);
  const double eq_rad2 = r2 / (r1 + r2);
  const mat3 mat0Trans = mtrans(&mat0);
  const v3 force = mmul(&mat0Trans, vadd(vadd(Fr, Fst), Fsr));
  const v3 moment = mmul(&mat0Trans, vadd(vadd(vadd(Mbr, Mtr), vmul(vmul(Msr, eq_rad1), rbpNorm)),
                                          vmul(vmul(Mst, eq_rad1), rbpNorm)));
  const v3 moment2 = mmul(&mat0Trans, vsub(vsub(vadd(vmul(vmul(Msr, eq_rad2), rbpNorm),
                                                     vmul(vmul(Mst, eq_rad2), rbpNorm)),
                                                Mbr),
                                           Mtr));
  o->shForce = vnorm(vadd(Fst, Fsr));
  o->tMoment = vnorm(Mtr);
  o->bMoment = vnorm(Mbr);
  const double r0 = r1 + r2;
  const v3 D = vsub(pos1, pos2);
  const double dist = sqrt(vdot(D, D));
  o->nForce = ((dist < r0) ? -1.0 : 1.0) * vnorm(Fr);
  o->cpos = vadd(pos2, vmul(D, r2 / (r1 + r2)));
  o->f1 = force;
  o->f2 = vmul(force, -1.0);
  o->m1 = moment;
  o->m2 = moment2;
}

/* CRotBondedInteraction constructor scaling (RotBondedInteraction.cpp:122-161);
 * the ARotBondedInteraction constructor values it overwrites are not kept. */
static void bond_constants(v3 pos1, double r1, v3 pos2, double r2, const esp_rot_bond_prms* prm,
                           const double* maxes, int do2d, obond* b)
{
  b->D = vsub(pos2, pos1);
  b->dist = sqrt(vdot(b->D, b->D));
  double effR = 1.0, effL = 1.0, effA = 1.0, momentJ = 1.0, momentI = 1.0;
  if (prm->scaling) {
    if (!do2d) {
      if (prm->meanR_scaling) effR = 0.5 * b->dist;
      else effR = fmin(r1, r2);
      effL = b->dist;
      effA = effR * effR;
      momentJ = 0.5 * effR * effR * effR * effR;
      momentI = 0.5 * momentJ;
    }
  }
  b->kr = prm->kr * effA / effL;
  b->ks = prm->ks * effA / effL;
  b->kb = prm->kb * momentI / effL;
  b->kt = prm->kt * momentJ / effL;
  b->max_n = maxes[0] * effA;
  b->max_s = maxes[1] * effA;
  b->max_b = maxes[3] * momentI / effR;
  b->max_t = maxes[2] * momentJ / effR;
  b->nForce = b->shForce = b->tMoment = b->bMoment = 0.0;
}

/* CRotBondedInteraction::broken (RotBondedInteraction.cpp:175-219). */
static int bond_broken(const obond* b, const esp_rot_bond_prms* prm)
{
  double shearStress = b->shForce / b->max_s + prm->beta2 * b->tMoment / b->max_t;
  double normalStress = b->nForce / b->max_n + prm->beta1 * b->bMoment / b->max_b;
  return (normalStress > prm->truncated) || (shearStress > (1.0 - normalStress));
}

typedef struct { v3 f1, m1, f2, m2, cpos; int touching; } fric_out;

/* CRotFrictionInteraction::calcSimpleForces (RotFricInteraction.cpp:171-231).
 * The elastic part is returned separately (fn) because the reference applies it
 * with its own applyForce calls before the frictional part. */
static void rot_friction_law(v3 pos1, v3 vel1, v3 w1, double r1, v3 pos2, v3 vel2, v3 w2,
                             double r2, double k, double ks, double mu_d, double mu_s, double dt,
                             v3* Ffric, int* slipping, double* ediss, v3* fn, fric_out* o)
{
  double mu_current = *slipping ? mu_d : mu_s;
  const v3 D = vsub(pos1, pos2);
  double dist = vdot(D, D);
  const double eq_dist = r1 + r2;
  if (dist < (eq_dist * eq_dist)) {
    dist = sqrt(dist);
    const v3 force = vmul(D, k * (dist - eq_dist) / dist);
    const v3 pos = vadd(pos2, vmul(D, r2 / eq_dist));
    *fn = force;
    const v3 vp1 = vadd(vel1, vcross(w1, vsub(pos, pos1)));
    const v3 vp2 = vadd(vel2, vcross(w2, vsub(pos, pos2)));
    v3 ds = vmul(vsub(vp2, vp1), dt);
    ds = vsub(ds, vmul(D, vdot(ds, D) / vnorm2(D)));
    *Ffric = vadd(*Ffric, vmul(ds, ks));
    const double FfricNorm = vnorm(*Ffric);
    const double forceNorm = vnorm(force);
    if (FfricNorm > forceNorm * mu_current) {
      *Ffric = vmul(*Ffric, (mu_d * forceNorm) / FfricNorm);
      *slipping = 1;
      *ediss = fabs(vdot(*Ffric, ds));
    } else {
      *slipping = 0;
      *ediss = 0.0;
    }
    o->m1 = vcross(vsub(pos, pos1), *Ffric);
    o->m2 = vcross(vsub(pos, pos2), vneg(*Ffric));
    o->f1 = *Ffric;
    o->f2 = vmul(*Ffric, -1.0);
    o->cpos = pos;
    o->touching = 1;
  } else {
    *Ffric = V(0.0, 0.0, 0.0);
    *slipping = 0;
    *ediss = 0.0;
    o->touching = 0;
  }
}

/* CRotFrictionInteraction::calcRigidBodyForces (RotFricInteraction.cpp:237-343),
 * the rigid=True form: the slip increment takes the tangential part of the
 * translational velocity difference plus the rotational one, and a stored
 * shear force of a touching, sticking pair first turns with the pair as a
 * rigid body.  touching is the interaction's flag from its last step.  Outputs
 * as rot_friction_law. */
static void rot_friction_rigid_law(v3 pos1, v3 vel1, v3 w1, double r1, v3 pos2, v3 vel2, v3 w2,
                                   double r2, double k, double ks, double mu_d, double mu_s,
                                   double dt, int touching, v3* Ffric, int* slipping,
                                   double* ediss, v3* fn, fric_out* o)
{
  const v3 D = vsub(pos1, pos2);
  double dist = vdot(D, D);
  const double eq_dist = r1 + r2;
  if (dist < (eq_dist * eq_dist)) {
    dist = sqrt(dist);
    const v3 force = vmul(D, k * (dist - eq_dist) / dist);
    const v3 pos = vadd(pos2, vmul(D, r2 / eq_dist));
    *fn = force;
    const v3 vp1_rot = vcross(w1, vsub(pos, pos1));
    const v3 vp2_rot = vcross(w2, vsub(pos, pos2));
    const v3 dv_trs = vsub(vel2, vel1);
    const v3 dv_trs_s = vsub(dv_trs, vmul(D, vdot(dv_trs, D) / vnorm2(D)));
    const v3 ds = vmul(vadd(vsub(vp2_rot, vp1_rot), dv_trs_s), dt);
    const double forceNorm = vnorm(force);
    if (!*slipping) {
      if (!touching) {
        *Ffric = V(0.0, 0.0, 0.0); /* first touch */
      } else {
        /* the pair's rigid-body rotation: spin about the axis plus the
         * rotation of the axis itself */
        const v3 rbp = vsub(pos2, pos1);
        const v3 vbp = vsub(vel2, vel1);
        const double rbp0 = vnorm(rbp);
        const v3 omiga_s = vmul(vadd(w1, w2), 0.5);
        const v3 omiga_spin = vdiv(vmul(rbp, vdot(omiga_s, rbp)), rbp0 * rbp0);
        const v3 omiga_m = vdiv(vcross(rbp, vbp), rbp0 * rbp0);
        const v3 d_Ffric = vmul(vcross(vadd(omiga_spin, omiga_m), *Ffric), dt);
        *Ffric = vadd(*Ffric, d_Ffric);
      }
      if (vnorm(vadd(*Ffric, vmul(ds, ks))) > forceNorm * mu_s) {
        *slipping = 1;
        *Ffric = vdiv(vmul(ds, mu_d * forceNorm), vnorm(ds));
        *ediss = fabs(vdot(*Ffric, ds));
      } else {
        *Ffric = vadd(*Ffric, vmul(ds, ks));
        *ediss = 0.0;
      }
    } else {
      if (vnorm(ds) > 1.0e-8) {
        *Ffric = vdiv(vmul(ds, mu_d * forceNorm), vnorm(ds));
        *ediss = fabs(vdot(*Ffric, ds));
      } else {
        *slipping = 0;
        *ediss = 0.0;
      }
    }
    o->m1 = vcross(vsub(pos, pos1), *Ffric);
    o->m2 = vcross(vsub(pos, pos2), vneg(*Ffric));
    o->f1 = *Ffric;
    o->f2 = vmul(*Ffric, -1.0);
    o->cpos = pos;
    o->touching = 1;
  } else {
    *Ffric = V(0.0, 0.0, 0.0);
    *slipping = 0;
    *ediss = 0.0;
    o->touching = 0;
  }
}

/* Scaled stiffnesses of the pair laws (constructors). */
static double elastic_k(double k, int scaling, double r1, double r2, int do2d)
{ /* ElasticInteraction.cpp:25-46 */
  double effR = 1.0, effA = 1.0, effL = 1.0;
  double r0 = r1 + r2;
  if (scaling) {
    if (!do2d) effR = 0.5 * r0;
    effL = r0;
    effA = effR * effR;
  }
  return k * effA / effL;
}
static double rotelastic_k(double kr, int scaling, double r1, double r2, int do2d)
{ /* RotElasticInteraction.cpp:40-57 */
  double f = 1.0;
  if (scaling) {
    if (!do2d) f = 0.5 * (r1 + r2);
  }
  return f * kr;
}
static void friction_k(double k, double ks, int scaling, int meanR, double r1, double r2,
                       int do2d, double* ko, double* kso)
{ /* RotFricInteraction.cpp:103-146 */
  double effR = 1.0, effL = 1.0, effA = 1.0;
  double r0 = r1 + r2;
  if (scaling) {
    if (!do2d) {
      if (meanR) effR = 0.5 * r0;
      else effR = fmin(r1, r2);
      effL = r0;
      effA = effR * effR;
    }
  }
  *ko = k * effA / effL;
  *kso = ks * effA / effL;
}

/* ------------------------------------------------------------------------- */
/* Groups                                                                      */
/* ------------------------------------------------------------------------- */
static int name_taken(const eso_engine* e, const char* name)
{
  for (int i = 0; i < e->ndg; i++) if (!strcmp(e->dg[i].name, name)) return 1;
  for (int i = 0; i < e->nbg; i++) if (!strcmp(e->bg[i].name, name)) return 1;
  for (int i = 0; i < e->nwg; i++) if (!strcmp(e->wg[i].name, name)) return 1;
  for (int i = 0; i < e->ngv; i++) if (!strcmp(e->gv[i].name, name)) return 1;
  return 0;
}
static int find_dg(const eso_engine* e, const char* n)
{
  for (int i = 0; i < e->ndg; i++) if (!strcmp(e->dg[i].name, n)) return i;
  return -1;
}
static int find_bg(const eso_engine* e, const char* n)
{
  for (int i = 0; i < e->nbg; i++) if (!strcmp(e->bg[i].name, n)) return i;
  return -1;
}
static int find_wall(const eso_engine* e, const char* n)
{
  for (int i = 0; i < e->nwl; i++) if (!strcmp(e->wl[i].name, n)) return i;
  return -1;
}

/* std::map<string,...> iteration order: names ascending. */
static int* name_order(int n, const char* (*get)(const eso_engine*, int), const eso_engine* e)
{
  int* o = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) o[i] = i;
  for (int i = 1; i < n; i++) {
    int v = o[i], j = i - 1;
    while (j >= 0 && strcmp(get(e, o[j]), get(e, v)) > 0) { o[j + 1] = o[j]; j--; }
    o[j + 1] = v;
  }
  return o;
}
static const char* dg_name(const eso_engine* e, int i) { return e->dg[i].name; }
static const char* bg_name(const eso_engine* e, int i) { return e->bg[i].name; }
static const char* wg_name(const eso_engine* e, int i) { return e->wg[i].name; }
static const char* gv_name(const eso_engine* e, int i) { return e->gv[i].name; }
static const char* dp_key(const eso_engine* e, int i) { return e->dp[i].key; }

/* CDamping::calcForces / CRotDamping::calcForces (Damping.hpp:87-107,
 * RotDamping.hpp:58-84): v_rel = v - v_ref + s*dt*(f - v_rel*visc*m) iterated
 * from zero while error*error > s_limit2 (1e-12, error = norm2 of the change)
 * and count < maxit; reaching maxit gives zero.  v_ref = 0 (DampingPrms). */
static v3 damping_vrel(v3 v, v3 frc, double s, double mass, double visc, double dt, int maxit)
{
  v3 v_rel = V(0.0, 0.0, 0.0);
  double error = 1.0;
  int count = 0;
  while ((error * error > 1e-12) & (count < maxit)) {
    count++;
    v3 v_old = v_rel;
    double sdt = s * dt;
    v3 inner = vsub(frc, vmul(vmul(v_rel, visc), mass));
    v_rel = vadd(vsub(v, V(0.0, 0.0, 0.0)), V(sdt * inner.x, sdt * inner.y, sdt * inner.z));
    error = vnorm2(vsub(v_rel, v_old));
  }
  if (count >= maxit) v_rel = V(0.0, 0.0, 0.0);
  return v_rel;
}

/* CLocalDamping / CRotLocalDamping::calcForces (LocalDamping.hpp:82-117,
 * RotLocalDamping.hpp:82-117), one component */
static double local_damping_c(double f, double v)
{
  double d = fabs(f);
  if (v < 0) d *= -1.;
  if (v == 0) d = 0.0;
  return d;
}

static void bg_keys(obgroup* G)
{
  if (!G->keys_dirty) return;
  G->nkeys = 0;
  GROW(G->keys, G->n, G->capkeys);
  for (int i = 0; i < G->n; i++) G->keys[G->nkeys++] = pkey(G->v[i].id1, G->v[i].id2);
  qsort(G->keys, G->nkeys, sizeof(long long), cmp_ll);
  G->keys_dirty = 0;
}

/* PIS_E::isIn for a bonded (excluding) group (pi_storage_e.hpp:176-188). */
static int bg_isin(eso_engine* e, int g, int id1, int id2)
{
  obgroup* G = &e->bg[g];
  bg_keys(G);
  return bsearch_ll(G->keys, G->nkeys, pkey(id1, id2)) >= 0;
}

/* isExcluded (pi_storage_ne.hpp:31-43, pi_storage_ed.hpp:48-60). */
static int dg_excluded(eso_engine* e, const ogroup* G, int id1, int id2)
{
  for (int x = 0; x < G->nex; x++)
    if (bg_isin(e, G->ex[x], id1, id2)) return 1;
  return 0;
}

/* the tag test of PIS_ED_T / PIS_NE_T::update (pi_storage_ed_t.hpp:55-70):
 * the pair's smaller tag against (tag1, mask1), the larger against (tag2, mask2) */
static int dg_tags_match(const ogroup* G, int ta, int tb)
{
  if (!G->tagged) return 1;
  int t1 = ta, t2 = tb;
  if (t1 > t2) {
    int th = t1;
    t1 = t2;
    t2 = th;
  }
  return ((t1 & G->mask1) == (G->tag1 & G->mask1)) && ((t2 & G->mask2) == (G->tag2 & G->mask2));
}

static void dg_push(eso_engine* e, ogroup* G, int i1, int i2)
{
  GROW(G->v, G->n, G->cap);
  odyn* d = &G->v[G->n++];
  memset(d, 0, sizeof *d);
  d->i1 = i1;
  d->i2 = i2;
  d->id1 = e->P[i1].id;
  d->id2 = e->P[i2].id;
  double r1 = e->P[i1].rad, r2 = e->P[i2].rad;
  if (G->kind == G_NE_ELASTIC) d->k = elastic_k(G->k, G->scaling, r1, r2, e->do2d);
  else if (G->kind == G_NE_ROTELASTIC) d->k = rotelastic_k(G->k, G->scaling, r1, r2, e->do2d);
  else friction_k(G->k, G->ks, G->scaling, G->meanR_scaling, r1, r2, e->do2d, &d->k, &d->ks);
}

static void store_full_list(eso_engine* e, const pairlist* L)
{
  e->nlast = 0;
  GROW(e->last_full, L->n, e->caplast);
  for (int i = 0; i < L->n; i++)
    e->last_full[e->nlast++] = pkey(e->P[L->a[i]].id, e->P[L->b[i]].id);
  qsort(e->last_full, e->nlast, sizeof(long long), cmp_ll);
}

/* PIS_NE::update (pi_storage_ne.hpp:49-102) and PIS_ED::update (pi_storage_ed.hpp:84-161). */
static void dg_update(eso_engine* e, ogroup* G)
{
  pairlist L = {0};
  if (G->update_ts != e->timestamp) {
    if (G->kind == G_ED_ROTFRICTION) {
      int w = 0;
      for (int i = 0; i < G->n; i++) {
        odyn d = G->v[i];
        /* isPersistent (RotFricInteraction.cpp:345-351) */
        v3 D = vsub(e->P[d.i1].pos, x_image(e, e->P[d.i1].pos, e->P[d.i2].pos));
        double dist = vdot(D, D);
        double eq = e->P[d.i1].rad + e->P[d.i2].rad;
        if (dist <= (eq * eq)) G->v[w++] = d;
      }
      G->n = w;
    } else {
      G->n = 0;
    }
    nt_pairs(e, &L, 0);
  } else {
    nt_pairs(e, &L, 1);
  }
  long long* present = NULL;
  int npresent = 0;
  if (G->kind == G_ED_ROTFRICTION) {
    present = malloc(sizeof(long long) * (size_t)(G->n > 0 ? G->n : 1));
    for (int i = 0; i < G->n; i++) present[npresent++] = pkey(G->v[i].id1, G->v[i].id2);
    qsort(present, npresent, sizeof(long long), cmp_ll);
  }
  for (int p = 0; p < L.n; p++) {
    int i1 = L.a[p], i2 = L.b[p];
    int id1 = e->P[i1].id, id2 = e->P[i2].id;
    if (!dg_tags_match(G, e->P[i1].tag, e->P[i2].tag)) continue;
    if (dg_excluded(e, G, id1, id2)) continue;
    if (G->kind == G_ED_ROTFRICTION && bsearch_ll(present, npresent, pkey(id1, id2)) >= 0)
      continue;
    dg_push(e, G, i1, i2);
  }
  free(present);
  free(L.a);
  free(L.b);
  G->update_ts = e->timestamp;
}

/* Pair-event log of a tracked group, after a step's interaction update: a
 * pair key whose stored count differs from the previous step's gives one event
 * per unit of difference, kind +1 (an interaction created) or -1 (removed).
 * Keys compare as (id1, id2), lower id first, so events come out in that order. */
static void log_pair_events(eso_engine* e, ogroup* G)
{
  long long* cur = malloc(sizeof(long long) * (size_t)(G->n > 0 ? G->n : 1));
  for (int i = 0; i < G->n; i++) cur[i] = pkey(G->v[i].id1, G->v[i].id2);
  qsort(cur, (size_t)G->n, sizeof(long long), cmp_ll);
  int a = 0, b = 0;
  while (a < G->nevprev || b < G->n) {
    long long k;
    int kind;
    if (b == G->n || (a < G->nevprev && G->evprev[a] < cur[b])) {
      k = G->evprev[a++];
      kind = -1;
    } else if (a == G->nevprev || cur[b] < G->evprev[a]) {
      k = cur[b++];
      kind = 1;
    } else {
      a++;
      b++;
      continue;
    }
    GROW(G->ev, G->nev + 3, G->capev);
    G->ev[G->nev++] = (int)e->t;
    G->ev[G->nev++] = (int)(k >> 32);
    G->ev[G->nev++] = (int)(k & 0xffffffffLL);
    G->ev[G->nev++] = kind;
  }
  free(G->evprev);
  G->evprev = cur;
  G->nevprev = G->n;
}

/* CEWallInteractionGroup::Update (EWallInteractionGroup.hpp:57-82) with
 * getParticlesAtPlane (ntable.hpp:674-690) and the CElasticWallInteraction
 * constructor (EWallInteraction.hpp:26-36). */
static void wg_update(eso_engine* e, owgroup* G)
{
  owall* W = &e->wl[G->wall];
  G->n = 0;
  for (int i = 0; i < e->np; i++) {
    double dist = vdot(vsub(e->P[i].pos, W->origin), W->normal);
    /* the tagged group keeps only candidates with the right tag (TaggedEWallInteractionGroup.hpp:42-44) */
    if (G->tagged_only && (e->P[i].tag & G->mask) != (G->tag & G->mask)) continue;
    if (fabs(dist) < e->dim) {
      GROW(G->pidx, G->n, G->cap);
      G->inner = realloc(G->inner, sizeof(int) * (size_t)G->cap);
      G->keff = realloc(G->keff, sizeof(double) * (size_t)G->cap);
      G->tagged = realloc(G->tagged, sizeof(int) * (size_t)G->cap);
      double f = 1.0;
      if (!e->do2d) f *= 3.141592654 * e->P[i].rad;
      G->pidx[G->n] = i;
      G->inner[G->n] = nt_inner(e, e->P[i].pos);
      G->keff[G->n] = f * G->k;
      G->tagged[G->n] = G->bonded && ((e->P[i].tag & G->mask) == (G->tag & G->mask));
      G->n++;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Step                                                                        */
/* ------------------------------------------------------------------------- */
static void log_rebuild(eso_engine* e, long long step)
{
  GROW(e->rebuild_steps, e->nreb, e->capreb);
  e->rebuild_steps[e->nreb++] = (int)step;
}

/* TSubLattice::searchNeighbors (SubLattice.hpp:2395-2410): rebuild the particle
 * array (timestamp++), rebuild interactions, reset displacements. */
static void search_neighbours(eso_engine* e, long long step)
{
  nt_build(e);
  e->timestamp++;
  for (int i = 0; i < e->np; i++) e->P[i].oldpos = e->P[i].pos;
  /* keep the full list of this rebuild for the neighbour-set queries */
  pairlist L = {0};
  nt_pairs(e, &L, 0);
  store_full_list(e, &L);
  free(L.a);
  free(L.b);
  log_rebuild(e, step);
}

/* TSubLattice::checkNeighbors (SubLattice.hpp:2473-2519). */
static int check_neighbours(eso_engine* e)
{
  double mdsqr = 0;
  double alpha = 0.5 * e->alpha;
  double srsqr = alpha * alpha;
  int i = 0;
  while (i < e->np && mdsqr < srsqr) {
    v3 d = vsub(e->P[i].pos, e->P[i].oldpos);
    double sqdisp = vdot(d, d);
    mdsqr = ((mdsqr < sqdisp) ? sqdisp : mdsqr);
    i++;
  }
  return mdsqr > srsqr;
}

/* TSubLattice::updateInteractions (SubLattice.hpp:2416-2466). */
/* CParticle::sigma (Particle.cpp:405-408, m_div_vol from Particle.cpp:52). */
static mat3 particle_sigma(const opart* p)
{
  const double div_vol = 1.0 / (4.0 / 3.0 * M_PI * p->rad * p->rad * p->rad);
  mat3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = p->sigma[i][j] * div_vol;
  return r;
}

/* Matrix3::dot (Matrix3.cpp:320-333): res(i,j) += a(i,k) b(k,j) from zero. */
static mat3 mdot(const mat3* a, const mat3* b)
{
  mat3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      r.m[i][j] = 0.0;
      for (int k = 0; k < 3; k++) r.m[i][j] += a->m[i][k] * b->m[k][j];
    }
  return r;
}

/* BrittleBeamSCInteraction::broken (BrittleBeamSC.cpp:103-148): truncated
 * Mohr-Coulomb test on the two particles' mean stress in the bond frame. */
static int sc_broken(const eso_engine* e, const obgroup* G, const obond* b)
{
  const opart* p1 = &e->P[b->i1];
  const opart* p2 = &e->P[b->i2];
  v3 d = vsub(x_image(e, p1->pos, p2->pos), p1->pos);
  v3 n = vdiv(d, vnorm(d));
  v3 s;
  if (n.x != 1.0 || n.y != 0.0 || n.z != 0.0) {
    v3 c = vcross(n, V(1.0, 0.0, 0.0));
    s = vdiv(c, vnorm(c));
  } else {
    v3 c = vcross(n, V(0.0, 1.0, 0.0));
    s = vdiv(c, vnorm(c));
  }
  v3 t = vcross(n, s);
  /* Matrix3(n,s,t) holds the vectors as columns; its transpose as rows */
  mat3 tm;
  tm.m[0][0] = n.x; tm.m[0][1] = n.y; tm.m[0][2] = n.z;
  tm.m[1][0] = s.x; tm.m[1][1] = s.y; tm.m[1][2] = s.z;
  tm.m[2][0] = t.x; tm.m[2][1] = t.y; tm.m[2][2] = t.z;
  mat3 s1 = particle_sigma(p1), s2 = particle_sigma(p2), avg, sym;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) avg.m[i][j] = 0.5 * (s1.m[i][j] + s2.m[i][j]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sym.m[i][j] = 0.5 * (avg.m[i][j] + avg.m[j][i]);
  mat3 tt = mtrans(&tm);
  mat3 a = mdot(&tm, &sym);
  mat3 rot = mdot(&a, &tt);
  double sigma_n = rot.m[0][0];
  double tau_x = rot.m[0][1];
  double tau_y = rot.m[0][2];
  double tau = sqrt(tau_x * tau_x + tau_y * tau_y);
  return (sigma_n < -1.0 * G->sc_sigma3) || (tau > G->sc_cohesion + sigma_n * G->sc_tan);
}

static void update_interactions(eso_engine* e)
{
  int need_update = 0;
  int* bo = name_order(e->nbg, bg_name, e);
  for (int k = 0; k < e->nbg; k++) {
    obgroup* G = &e->bg[bo[k]];
    int w = 0, res = 0;
    for (int i = 0; i < G->n; i++) {
      obond b = G->v[i];
      if (G->sc ? sc_broken(e, G, &b) : bond_broken(&b, &G->prm)) {
        res = 1;
        e->P[b.i1].flag = 1;
        e->P[b.i2].flag = 1;
        GROW(e->brk, e->nbrk + 2, e->capbrk);
        e->brk[e->nbrk++] = (int)e->t;
        e->brk[e->nbrk++] = b.id1;
        e->brk[e->nbrk++] = b.id2;
        G->keys_dirty = 1;
      } else {
        G->v[w++] = b;
      }
    }
    G->n = w;
    need_update = need_update || res;
  }
  free(bo);
  if ((e->timestamp > e->last_ns) || need_update) {
    e->num_host_phases++;
    int* dord = name_order(e->ndg, dg_name, e);
    for (int k = 0; k < e->ndg; k++) dg_update(e, &e->dg[dord[k]]);
    free(dord);
    int* word = name_order(e->nwg, wg_name, e);
    for (int k = 0; k < e->nwg; k++) wg_update(e, &e->wg[word[k]]);
    free(word);
    e->last_ns = e->timestamp;
  }
}

/* TSubLattice::zeroForces + calcForces + integrate (SubLattice.hpp:2066-2218). */
static void calc_and_integrate(eso_engine* e)
{
  for (int i = 0; i < e->np; i++) {
    opart* p = &e->P[i];
    p->force = V(0.0, 0.0, 0.0);
    p->moment = V(0.0, 0.0, 0.0);
    p->fabs_sum = 0.0;
    p->mabs_sum = 0.0;
    memset(p->sigma, 0, sizeof p->sigma);
  }
  for (int w = 0; w < e->nwl; w++) e->wl[w].force = V(0.0, 0.0, 0.0);

  /* walls (m_WIG, name order) */
  int* word = name_order(e->nwg, wg_name, e);
  for (int k = 0; k < e->nwg; k++) {
    owgroup* G = &e->wg[word[k]];
    owall* W = &e->wl[G->wall];
    /* CBWallInteractionGroup::calcForces: bonded interactions first
     * (CBondedWallInteraction::calcForces, BWallInteraction.hpp:35-52) */
    for (int c = 0; G->bonded && c < G->n; c++) {
      if (!G->tagged[c]) continue;
      opart* p = &e->P[G->pidx[c]];
      double scale = e->do2d ? 1.0 : 3.1415926536 * p->rad;
      double kb = G->k * scale;
      v3 D = vsub(vsub(p->pos, p->initpos), vsub(W->origin, W->oldpos));
      v3 force = vmul(D, kb);
      apply_force(p, vmul(force, -1.0), p->pos);
      if (G->inner[c]) W->force = vsub(W->force, force);
    }
    for (int c = 0; c < G->n; c++) {
      if (G->bonded && G->tagged[c]) continue;
      opart* p = &e->P[G->pidx[c]];
      double dist = vdot(vsub(p->pos, W->origin), W->normal);
      if (dist < p->rad) {
        v3 force = vmul(W->normal, G->keff[c] * (p->rad - dist));
        v3 pos = vsub(p->pos, vmul(W->normal, dist));
        apply_force(p, force, pos);
        if (G->inner[c]) W->force = vsub(W->force, vmul(force, -1.0));
      }
    }
  }
  free(word);

  /* single-particle interaction groups: gravity (name order) */
  int* gord = name_order(e->ngv, gv_name, e);
  for (int k = 0; k < e->ngv; k++) {
    v3 g = e->gv[gord[k]].g;
    for (int i = 0; i < e->np; i++) apply_force(&e->P[i], vmul(g, e->P[i].mass), e->P[i].pos);
  }
  free(gord);

  /* dynamic groups (m_dpis, name order), each in list order */
  int* dord = name_order(e->ndg, dg_name, e);
  for (int k = 0; k < e->ndg; k++) {
    ogroup* G = &e->dg[dord[k]];
    for (int c = 0; c < G->n; c++) {
      odyn* d = &G->v[c];
      opart* p1 = &e->P[d->i1];
      opart* p2 = &e->P[d->i2];
      const v3 x2 = x_image(e, p1->pos, p2->pos);
      if (G->kind == G_NE_ELASTIC) {
        /* CElasticInteraction::calcForces (ElasticInteraction.cpp:54-71) */
        v3 D = vsub(p1->pos, x2);
        double dist = vdot(D, D);
        double eq_dist = p1->rad + p2->rad;
        if (dist < (eq_dist * eq_dist)) {
          dist = sqrt(dist);
          v3 force = vmul(D, d->k * (dist - eq_dist) / dist);
          v3 pos = vadd(x2, vmul(D, p2->rad / eq_dist));
          apply_force_at(p2, force, pos, x2);
          apply_force(p1, vmul(force, -1.0), pos);
          d->touching = 1;
        } else {
          d->touching = 0;
        }
      } else if (G->kind == G_NE_ROTELASTIC) {
        /* CRotElasticInteraction::calcForces (RotElasticInteraction.cpp:69-88) */
        v3 rb = vsub(p1->pos, x2);
        double rbNorm = vnorm(rb);
        double r_0Norm = p1->rad + p2->rad;
        v3 delta_r = vdiv(vmul(rb, rbNorm - r_0Norm), rbNorm);
        if (rbNorm < r_0Norm) {
          v3 force = vmul(delta_r, d->k);
          v3 pos = vadd(x2, vmul(rb, p2->rad / (p1->rad + p2->rad)));
          apply_force(p1, vmul(force, -1.0), pos);
          apply_force_at(p2, force, pos, x2);
        }
      } else {
        fric_out o;
        v3 fn = V(0, 0, 0);
        if (G->rigid)
          rot_friction_rigid_law(p1->pos, p1->vel, p1->angvel, p1->rad, x2, p2->vel, p2->angvel,
                                 p2->rad, d->k, d->ks, G->mu_d, G->mu_s, e->dt, d->touching,
                                 &d->ffric, &d->slipping, &d->ediss, &fn, &o);
        else
          rot_friction_law(p1->pos, p1->vel, p1->angvel, p1->rad, x2, p2->vel, p2->angvel,
                           p2->rad, d->k, d->ks, G->mu_d, G->mu_s, e->dt, &d->ffric, &d->slipping,
                           &d->ediss, &fn, &o);
        /* per particle the rigid form adds the friction force's moment before
         * the explicit moment (RotFricInteraction.cpp:324-327); both are
         * cross(pos - x, Ffric), so the sums agree with this order */
        if (o.touching) {
          apply_force_at(p2, fn, o.cpos, x2);
          apply_force(p1, vmul(fn, -1.0), o.cpos);
          apply_moment(p1, o.m1);
          apply_moment(p2, o.m2);
          apply_force(p1, o.f1, o.cpos);
          apply_force_at(p2, o.f2, o.cpos, x2);
          d->touching = 1;
        } else {
          d->touching = 0;
        }
      }
    }
  }
  free(dord);

  /* bonded groups (m_bpis, name order), each in list order */
  int* bo = name_order(e->nbg, bg_name, e);
  for (int k = 0; k < e->nbg; k++) {
    obgroup* G = &e->bg[bo[k]];
    for (int c = 0; c < G->n; c++) {
      obond* b = &G->v[c];
      opart* p1 = &e->P[b->i1];
      opart* p2 = &e->P[b->i2];
      bond_out o;
      const v3 x2 = x_image(e, p1->pos, p2->pos);
      rot_bond_law(p1->pos, p1->q, p1->rad, x2, p2->q, p2->rad, b->kr, b->ks, b->kt, b->kb,
                   b->D, &o);
      b->nForce = o.nForce;
      b->shForce = o.shForce;
      b->tMoment = o.tMoment;
      b->bMoment = o.bMoment;
      apply_force_at(p2, o.f2, o.cpos, x2);
      apply_force(p1, o.f1, o.cpos);
      apply_moment(p1, o.m1);
      apply_moment(p2, o.m2);
    }
  }
  free(bo);

  /* last, but not least - damping (m_damping, key order, SubLattice.hpp:2143-2147) */
  int* dpo = name_order(e->ndp, dp_key, e);
  for (int k = 0; k < e->ndp; k++) {
    const odamp* D = &e->dp[dpo[k]];
    for (int i = 0; i < e->np; i++) {
      opart* p = &e->P[i];
      if (D->type == ESP_DAMPING) {
        v3 vr = damping_vrel(p->vel, p->force, p->inv_mass, p->mass, D->visc, e->dt, D->maxit);
        double c = -1.0 * D->visc;
        v3 f = vmul(V(c * vr.x, c * vr.y, c * vr.z), p->mass);
        apply_force(p, f, p->pos);
      } else if (D->type == ESP_ROT_DAMPING) {
        double s = 1.0 / p->inert, in = p->inert;
        v3 vr = damping_vrel(p->angvel, p->moment, s, in, D->visc, e->dt, D->maxit);
        double c = -1.0 * D->visc;
        apply_moment(p, vmul(V(c * vr.x, c * vr.y, c * vr.z), in));
      } else {
        int rd = D->type == ESP_ROT_LOCAL_DAMPING;
        v3 v = rd ? p->angvel : p->vel, f = rd ? p->moment : p->force;
        v3 d = V(local_damping_c(f.x, v.x), local_damping_c(f.y, v.y), local_damping_c(f.z, v.z));
        double c = -1.0 * D->visc;
        v3 fd = V(c * d.x, c * d.y, c * d.z);
        if (rd) apply_moment(p, fd);
        else apply_force(p, fd, p->pos);
      }
    }
  }
  free(dpo);

  for (int i = 0; i < e->np; i++) integrate_particle(&e->P[i], e->dt, e->rot, e->do2d);
  for (int w = 0; w < e->nwl; w++) {
    owall* W = &e->wl[w];
    GROW(W->hist, W->nhist + 3, W->caphist);
    W->hist[W->nhist++] = W->force.x;
    W->hist[W->nhist++] = W->force.y;
    W->hist[W->nhist++] = W->force.z;
  }
}

/* CLatticeMaster::runOneStep (LatticeMaster.cpp:2602-2725), one worker. */
static void run_one_step(eso_engine* e)
{
  if (e->first_time) {
    search_neighbours(e, e->t + 1);
    e->first_time = 0;
  }
  e->t++;
  /* pre-time-step runnables: registered wall motions */
  for (int w = 0; w < e->nwl; w++)
    if (e->wl[w].has_motion) e->wl[w].origin = vadd(e->wl[w].origin, e->wl[w].step_motion);
  if (check_neighbours(e)) search_neighbours(e, e->t);
  update_interactions(e);
  for (int g = 0; g < e->ndg; g++)
    if (e->dg[g].track) log_pair_events(e, &e->dg[g]);
  calc_and_integrate(e);
}

/* ------------------------------------------------------------------------- */
/* C entry points                                                              */
/* ------------------------------------------------------------------------- */
const char* eso_create_error(void) { return g_create_err; }
const char* eso_last_error(const eso_engine* e) { return e ? e->err : g_create_err; }

int eso_create(const esp_config* cfg, eso_engine** out)
{
  g_create_err[0] = 0;
  if (!cfg || !out) { snprintf(g_create_err, sizeof g_create_err, "null argument"); return ESP_ERR_ARG; }
  if (cfg->particle_type != ESP_PTYPE_NROT && cfg->particle_type != ESP_PTYPE_ROT) {
    snprintf(g_create_err, sizeof g_create_err, "unknown particle type %d", cfg->particle_type);
    return ESP_ERR_ARG;
  }
  if (!(cfg->grid_spacing > 0) || !(cfg->verlet_dist >= 0)) {
    snprintf(g_create_err, sizeof g_create_err, "gridSpacing must be > 0 and verletDist >= 0");
    return ESP_ERR_ARG;
  }
  eso_engine* e = calloc(1, sizeof *e);
  e->cfg = *cfg;
  e->rot = cfg->particle_type == ESP_PTYPE_ROT;
  e->do2d = cfg->do_2d;
  e->dt = cfg->dt;
  e->dim = cfg->grid_spacing;
  e->alpha = cfg->verlet_dist;
  /* TSubLattice::initNeighborTable (SubLattice.hpp:455-478) */
  v3 mn = V(cfg->domain_min[0], cfg->domain_min[1], cfg->domain_min[2]);
  v3 mx = V(cfg->domain_max[0], cfg->domain_max[1], cfg->domain_max[2]);
  double nr = e->dim;
  double xs = mx.x - mn.x;
  if (cfg->circular_x) {
    /* TSubLattice::initNeighborTable(min, max, circ) (SubLattice.hpp:502-515):
     * the range is adjusted so that whole cells span the circular x size */
    if (fabs(xs / nr - lrint(xs / nr)) > 1e-6) nr = xs / floor(xs / nr);
    e->dim = nr;
  } else {
    xs = nr * ceil(xs / nr);
  }
  double ys = mx.y - mn.y; ys = nr * ceil(ys / nr);
  double zs = mx.z - mn.z; zs = nr * ceil(zs / nr);
  v3 grow = vsub(V(xs, ys, zs), vsub(mx, mn));
  v3 nmin = vsub(mn, vmul(grow, 0.5));
  v3 nmax = vadd(mx, vmul(grow, 0.5));
  /* ParallelParticleArray ctor (pp_array.hpp:61-103), 1x1x1 process grid */
  long nxg = lrint((nmax.x - nmin.x) / nr);
  long nyg = lrint((nmax.y - nmin.y) / nr);
  long nzg = lrint((nmax.z - nmin.z) / nr);
  if ((fabs((double)nxg - (nmax.x - nmin.x) / nr) > 1e-6) ||
      (fabs((double)nyg - (nmax.y - nmin.y) / nr) > 1e-6) ||
      (fabs((double)nzg - (nmax.z - nmin.z) / nr) > 1e-6) || nxg < 1 || nyg < 1 || nzg < 1) {
    snprintf(g_create_err, sizeof g_create_err, "size doesn't fit range");
    free(e);
    return ESP_ERR_ARG;
  }
  double ncells = (double)(nxg + 2) * (double)(nyg + 2) * (double)(nzg + 2);
  if (ncells > 2.0e8) {
    snprintf(g_create_err, sizeof g_create_err, "neighbour table too large (%.0f cells)", ncells);
    free(e);
    return ESP_ERR_CAPACITY;
  }
  e->gidx = e->gidy = e->gidz = -1;
  e->nx = (int)nxg + 2;
  if (cfg->circular_x) {
    if (nxg < 3) {
      snprintf(g_create_err, sizeof g_create_err, "a circular x dimension needs at least 3 cells");
      free(e);
      return ESP_ERR_ARG;
    }
    e->circ_x = 1;
    e->gidx = 0;
    e->nx = (int)nxg;
    e->period = nmax.x - nmin.x;
  }
  e->ny = (int)nyg + 2;
  e->nz = (int)nzg + 2;
  e->ncell = e->nx * e->ny * e->nz;
  e->p0 = nmin;
  e->first_time = 1;
  e->last_ns = -1;
  e->cell_start = calloc((size_t)e->ncell + 1, sizeof(int));
  e->cell_items = malloc(sizeof(int));
  *out = e;
  return ESP_OK;
}

void eso_destroy(eso_engine* e)
{
  if (!e) return;
  for (int i = 0; i < e->ndg; i++) { free(e->dg[i].v); free(e->dg[i].evprev); free(e->dg[i].ev); }
  for (int i = 0; i < e->nbg; i++) { free(e->bg[i].v); free(e->bg[i].keys); }
  for (int i = 0; i < e->nwl; i++) free(e->wl[i].hist);
  for (int i = 0; i < e->nwg; i++) { free(e->wg[i].pidx); free(e->wg[i].inner); free(e->wg[i].keff); }
  free(e->dg); free(e->bg); free(e->wl); free(e->wg); free(e->gv);
  free(e->P); free(e->idkey); free(e->conn); free(e->brk); free(e->rebuild_steps);
  free(e->last_full); free(e->cell_start); free(e->cell_items);
  free(e);
}

int eso_set_time_step_size(eso_engine* e, double dt)
{
  e->dt = dt;
  return ESP_OK;
}

int eso_add_particles(eso_engine* e, int n, const int* id, const int* tag, const double* pos,
                      const double* rad, const double* mass, const double* vel)
{
  if (n < 0 || (n > 0 && (!id || !pos || !rad))) return fail(e, ESP_ERR_ARG, "bad particle arrays");
  int np0 = e->np;
  for (int i = 0; i < n; i++) {
    if (id[i] < 0) { e->np = np0; return fail(e, ESP_ERR_ARG, "particle ids must be >= 0 (got %d)", id[i]); }
    opart p;
    memset(&p, 0, sizeof p);
    p.id = id[i];
    p.tag = tag ? tag[i] : 0;
    p.pos = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    p.rad = rad[i];
    double r = rad[i];
    p.mass = mass ? mass[i] : (e->do2d ? r * r : r * r * r);
    p.vel = vel ? V(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]) : V(0, 0, 0);
    p.oldpos = p.pos;
    p.initpos = p.pos;
    p.inv_mass = (p.mass != 0.0) ? 1.0 / p.mass : 0.0;
    p.is_dyn = 1;
    p.is_rot = e->rot;
    p.q.s = 1.0;
    if (e->rot) {
      p.inert = e->do2d ? 0.5 * p.mass * r * r : 0.4 * p.mass * r * r;
      p.inv_inert = 1.0 / p.inert;
    }
    /* NeighborTable::insert (ntable.hpp:231-255): outside -> not kept */
    if (nt_index(e, p.pos) == -1) continue;
    GROW(e->P, e->np, e->capp);
    e->P[e->np++] = p;
  }
  rebuild_idmap(e);
  for (int k = 1; k < e->nidkey; k++)
    if ((e->idkey[k] >> 32) == (e->idkey[k - 1] >> 32)) {
      int dup = (int)(e->idkey[k] >> 32);
      e->np = np0;
      rebuild_idmap(e);
      return fail(e, ESP_ERR_ARG, "duplicate particle id %d", dup);
    }
  return ESP_OK;
}

int eso_set_particle_density(eso_engine* e, int tag, int mask, double rho)
{
  const double pi = 3.141592654;
  for (int i = 0; i < e->np; i++) {
    opart* p = &e->P[i];
    if ((p->tag & mask) != (tag & mask)) continue;
    double r = p->rad;
    if (e->do2d) {
      p->mass = rho * pi * r * r;
      if (e->rot) p->inert = 0.5 * p->mass * r * r;
    } else {
      p->mass = (4.0 / 3.0) * rho * pi * r * r * r;
      if (e->rot) p->inert = 0.4 * p->mass * r * r;
    }
    if (e->rot) p->inv_inert = 1.0 / p->inert;
    p->inv_mass = (p->mass != 0.0) ? 1.0 / p->mass : 0.0;
  }
  return ESP_OK;
}

int eso_set_particle_non_dynamic(eso_engine* e, int tag)
{
  for (int i = 0; i < e->np; i++)
    if (e->P[i].tag == tag) { e->P[i].is_dyn = 0; e->P[i].is_rot = 0; }
  return ESP_OK;
}

int eso_set_particle_non_rotational(eso_engine* e, int tag)
{
  for (int i = 0; i < e->np; i++)
    if (e->P[i].tag == tag) e->P[i].is_rot = 0;
  return ESP_OK;
}

/* CLatticeMaster::setParticleVel / setParticleAngVel (LatticeMaster.cpp:1499) */
int eso_set_particle_vel(eso_engine* e, int id, const double v[3])
{
  for (int i = 0; i < e->np; i++)
    if (e->P[i].id == id) {
      e->P[i].vel = V(v[0], v[1], v[2]);
      return ESP_OK;
    }
  return fail(e, ESP_ERR_ARG, "unknown particle id %d", id);
}

int eso_set_particle_ang_vel(eso_engine* e, int id, const double w[3])
{
  if (!e->rot) return fail(e, ESP_ERR_ARG, "setParticleAngularVelocity needs rotational particles");
  for (int i = 0; i < e->np; i++)
    if (e->P[i].id == id) {
      e->P[i].angvel = V(w[0], w[1], w[2]);
      return ESP_OK;
    }
  return fail(e, ESP_ERR_ARG, "unknown particle id %d", id);
}

int eso_add_connections(eso_engine* e, int n, const int* id1, const int* id2, const int* tag)
{
  if (n < 0 || (n > 0 && (!id1 || !id2))) return fail(e, ESP_ERR_ARG, "bad connection arrays");
  for (int i = 0; i < n; i++) {
    GROW(e->conn, e->nconn + 3, e->capconn);
    e->conn[e->nconn++] = tag ? tag[i] : 0;
    e->conn[e->nconn++] = id1[i];
    e->conn[e->nconn++] = id2[i];
  }
  return ESP_OK;
}

static int add_bond_group(eso_engine* e, const char* name, const esp_rot_bond_prms* prm,
                          const double* per_bond_max, int n_per_bond,
                          const esp_brittle_beam_sc_prms* sc);

int eso_add_rot_bond_group(eso_engine* e, const char* name, const esp_rot_bond_prms* prm,
                           const double* per_bond_max, int n_per_bond)
{
  return add_bond_group(e, name, prm, per_bond_max, n_per_bond, NULL);
}

/* TRotSubLattice::addBrittleBeamSCIG (RotSubLattice.hpp:188-225): the bonds of
 * ARotBondedInteraction (scaled stiffnesses, ARotBondedInteraction.cpp:141-179)
 * with the stress criterion of BrittleBeamSCInteraction (BrittleBeamSC.cpp:73-79). */
int eso_add_brittle_beam_sc_group(eso_engine* e, const char* name,
                                  const esp_brittle_beam_sc_prms* sc)
{
  if (!sc) return fail(e, ESP_ERR_ARG, "bad group");
  esp_rot_bond_prms prm;
  memset(&prm, 0, sizeof prm);
  prm.kr = sc->kr;
  prm.ks = sc->ks;
  prm.kt = sc->kt;
  prm.kb = sc->kb;
  prm.max_n = prm.max_s = prm.max_t = prm.max_b = 1.0;
  prm.tag = sc->tag;
  prm.scaling = 1;
  prm.meanR_scaling = 1;
  prm.truncated = prm.beta1 = prm.beta2 = 1.0;
  return add_bond_group(e, name, &prm, NULL, 0, sc);
}

static int add_bond_group(eso_engine* e, const char* name, const esp_rot_bond_prms* prm,
                          const double* per_bond_max, int n_per_bond,
                          const esp_brittle_beam_sc_prms* sc)
{
  if (!e->rot) return fail(e, ESP_ERR_STATE, "rotational bonds need particle type RotSphere");
  if (!name || !prm || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad group");
  if (name_taken(e, name)) return fail(e, ESP_ERR_ARG, "interaction group '%s' exists", name);
  GROW(e->bg, e->nbg, e->capbg);
  obgroup* G = &e->bg[e->nbg];
  memset(G, 0, sizeof *G);
  strcpy(G->name, name);
  G->prm = *prm;
  G->keys_dirty = 1;
  if (sc) {
    G->sc = 1;
    G->sc_sigma3 = sc->tension_cutoff;
    G->sc_cohesion = sc->cohesion;
    G->sc_tan = tan(sc->friction_angle / 57.2957795);
  }
  /* connections of this tag, in arrival order; a repeated (sorted) id pair is
   * rejected by isIn at its second occurrence */
  int nc = 0;
  for (int c = 0; c < e->nconn; c += 3) nc += (e->conn[c] == prm->tag);
  keyidx* ck = malloc(sizeof(keyidx) * (size_t)(nc > 0 ? nc : 1));
  int* cpos = malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
  int k = 0;
  for (int c = 0; c < e->nconn; c += 3) {
    if (e->conn[c] != prm->tag) continue;
    int a = e->conn[c + 1], b = e->conn[c + 2];
    ck[k].k = pkey(a < b ? a : b, a < b ? b : a);
    ck[k].i = k;
    cpos[k] = c;
    k++;
  }
  if (per_bond_max && k != n_per_bond) {
    free(ck); free(cpos);
    return fail(e, ESP_ERR_ARG, "per_bond_max has %d entries for %d connections", n_per_bond, k);
  }
  unsigned char* first = calloc((size_t)(nc > 0 ? nc : 1), 1);
  sort_keyidx(ck, nc);
  for (int i = 0; i < nc; i++) first[ck[i].i] = (i == 0 || ck[i].k != ck[i - 1].k);
  for (int j = 0; j < nc; j++) {
    if (!first[j]) continue;
    int c = cpos[j];
    int a = e->conn[c + 1], b = e->conn[c + 2];
    int id1 = a < b ? a : b, id2 = a < b ? b : a;
    double maxes[4] = {prm->max_n, prm->max_s, prm->max_t, prm->max_b};
    if (per_bond_max) memcpy(maxes, per_bond_max + 4 * j, sizeof maxes);
    /* PIS_E::tryInsert (pi_storage_e.hpp:216-250) */
    int i1 = find_particle(e, id1), i2 = find_particle(e, id2);
    if (i1 < 0 || i2 < 0) continue;
    GROW(G->v, G->n, G->cap);
    obond* bd = &G->v[G->n++];
    memset(bd, 0, sizeof *bd);
    bd->id1 = id1;
    bd->id2 = id2;
    bd->i1 = i1;
    bd->i2 = i2;
    bond_constants(e->P[i1].pos, e->P[i1].rad, x_image(e, e->P[i1].pos, e->P[i2].pos),
                   e->P[i2].rad, prm, maxes, e->do2d, bd);
  }
  free(ck); free(cpos); free(first);
  e->nbg++;
  return ESP_OK;
}

static int add_dg(eso_engine* e, const char* name, int kind, double k, double ks, double mud,
                  double mus, int scaling, int meanR)
{
  if (!name || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad group name");
  if (name_taken(e, name)) return fail(e, ESP_ERR_ARG, "interaction group '%s' exists", name);
  GROW(e->dg, e->ndg, e->capdg);
  ogroup* G = &e->dg[e->ndg++];
  memset(G, 0, sizeof *G);
  strcpy(G->name, name);
  G->kind = kind;
  G->k = k;
  G->ks = ks;
  G->mu_d = mud;
  G->mu_s = mus;
  G->scaling = scaling;
  G->meanR_scaling = meanR;
  G->update_ts = 0;
  return ESP_OK;
}

int eso_add_nrot_elastic_group(eso_engine* e, const char* name, double k, int scaling)
{
  return add_dg(e, name, G_NE_ELASTIC, k, 0, 0, 0, scaling, 1);
}

int eso_add_rot_elastic_group(eso_engine* e, const char* name, double kr, int scaling)
{
  if (!e->rot) return fail(e, ESP_ERR_STATE, "RotElastic needs particle type RotSphere");
  return add_dg(e, name, G_NE_ROTELASTIC, kr, 0, 0, 0, scaling, 1);
}

int eso_add_rot_friction_group(eso_engine* e, const char* name, const esp_rot_friction_prms* p)
{
  if (!e->rot) return fail(e, ESP_ERR_STATE, "RotFriction needs particle type RotSphere");
  if (!p) return fail(e, ESP_ERR_ARG, "null parameters");
  int rc = add_dg(e, name, G_ED_ROTFRICTION, p->k_n, p->k_s, p->mu_d, p->mu_s, p->scaling,
                  p->meanR_scaling);
  if (rc == ESP_OK) e->dg[e->ndg - 1].rigid = p->rigid != 0;
  return rc;
}

int eso_add_exclusion(eso_engine* e, const char* excluding, const char* excluded)
{
  int b = find_bg(e, excluding);
  int d = find_dg(e, excluded);
  if (b < 0) return fail(e, ESP_ERR_ARG, "excluding group '%s' is not a bonded group", excluding);
  if (d < 0) return fail(e, ESP_ERR_ARG, "excluded group '%s' is not a dynamic group", excluded);
  ogroup* G = &e->dg[d];
  if (G->nex >= 16) return fail(e, ESP_ERR_CAPACITY, "too many exclusions");
  G->ex[G->nex++] = b;
  if (G->kind == G_ED_ROTFRICTION) { /* PIS_ED::addExIG purges (pi_storage_ed.hpp:20-40) */
    int w = 0;
    for (int i = 0; i < G->n; i++)
      if (!bg_isin(e, b, G->v[i].id1, G->v[i].id2)) G->v[w++] = G->v[i];
    G->n = w;
  }
  return ESP_OK;
}

int eso_add_wall(eso_engine* e, const char* name, const double o[3], const double n[3])
{
  if (!name || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad wall name");
  if (find_wall(e, name) >= 0) return fail(e, ESP_ERR_ARG, "wall '%s' exists", name);
  GROW(e->wl, e->nwl, e->capwl);
  owall* W = &e->wl[e->nwl++];
  memset(W, 0, sizeof *W);
  strcpy(W->name, name);
  W->origin = V(o[0], o[1], o[2]);
  W->oldpos = W->origin;
  W->normal = V(n[0], n[1], n[2]);
  return ESP_OK;
}

int eso_add_elastic_wall_group(eso_engine* e, const char* name, const char* wall, double k)
{
  int w = find_wall(e, wall);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", wall ? wall : "");
  if (!name || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad group name");
  if (name_taken(e, name)) return fail(e, ESP_ERR_ARG, "interaction group '%s' exists", name);
  GROW(e->wg, e->nwg, e->capwg);
  owgroup* G = &e->wg[e->nwg++];
  memset(G, 0, sizeof *G);
  strcpy(G->name, name);
  G->wall = w;
  G->k = k;
  return ESP_OK;
}

int eso_add_bonded_wall_group(eso_engine* e, const char* name, const char* wall, double k,
                              int tag, int mask)
{
  int rc = eso_add_elastic_wall_group(e, name, wall, k);
  if (rc) return rc;
  owgroup* G = &e->wg[e->nwg - 1];
  G->bonded = 1;
  G->tag = tag;
  G->mask = mask;
  return ESP_OK;
}

int eso_set_wall_group_tags(eso_engine* e, const char* name, int tag, int mask)
{
  for (int k = 0; k < e->nwg; k++)
    if (!strcmp(e->wg[k].name, name ? name : "")) {
      if (e->wg[k].bonded) return fail(e, ESP_ERR_ARG, "a bonded wall group selects its tags itself");
      e->wg[k].tagged_only = 1;
      e->wg[k].tag = tag;
      e->wg[k].mask = mask;
      return ESP_OK;
    }
  return fail(e, ESP_ERR_ARG, "not an elastic wall group");
}

int eso_add_gravity(eso_engine* e, const char* name, const double g[3])
{
  if (!name || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad group name");
  if (name_taken(e, name)) return fail(e, ESP_ERR_ARG, "interaction group '%s' exists", name);
  GROW(e->gv, e->ngv, e->capgv);
  ograv* G = &e->gv[e->ngv++];
  strcpy(G->name, name);
  G->g = V(g[0], g[1], g[2]);
  return ESP_OK;
}

int eso_add_damping_group(eso_engine* e, int type, const char* name, double viscosity,
                          int max_iterations)
{
  static const char* tname[4] = {"Damping", "RotDamping", "LocalDamping", "RotLocalDamping"};
  if (type < 0 || type > 3) return fail(e, ESP_ERR_ARG, "unknown damping type");
  if ((type == ESP_ROT_DAMPING || type == ESP_ROT_LOCAL_DAMPING) && !e->rot)
    return fail(e, ESP_ERR_ARG, "Trying to setup Damping of unknown type: %s", tname[type]);
  if (!name || strlen(name) >= 128) return fail(e, ESP_ERR_ARG, "bad group name");
  const char* key = (type == ESP_DAMPING || type == ESP_ROT_DAMPING) ? tname[type] : name;
  for (int k = 0; k < e->ndp; k++)
    if (!strcmp(e->dp[k].key, key)) return ESP_OK; /* map insert keeps the first */
  if (e->ndp >= 4) return fail(e, ESP_ERR_CAPACITY, "more than 4 damping groups");
  odamp* D = &e->dp[e->ndp++];
  strcpy(D->key, key);
  D->type = type;
  D->visc = viscosity;
  D->maxit = max_iterations;
  return ESP_OK;
}

int eso_move_wall_by(eso_engine* e, const char* name, const double d[3])
{
  int w = find_wall(e, name);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", name ? name : "");
  e->wl[w].origin = vadd(e->wl[w].origin, V(d[0], d[1], d[2]));
  return ESP_OK;
}

/* applyForceToWall(interactionName, force) (LsmMpiPy.cpp:1208 ->
 * TSubLattice::applyForceToWall, SubLattice.hpp:2935-2950): the wall of the
 * group moves along the force direction until the group's wall force balances
 * the force.  Elastic groups: CEWallInteractionGroup::applyForce
 * (EWallInteractionGroup.hpp:93-164); bonded groups:
 * CBWallInteractionGroup::applyForce (BWallInteractionGroup.hpp:71-145).
 * Candidates in their list order, sums from zero in that order. */
int eso_apply_force_to_wall(eso_engine* e, const char* name, const double Fa[3])
{
  int g = -1;
  for (int i = 0; i < e->nwg; i++)
    if (!strcmp(e->wg[i].name, name ? name : "")) g = i;
  if (g < 0) return fail(e, ESP_ERR_ARG, "unknown wall interaction group '%s'", name ? name : "");
  owgroup* G = &e->wg[g];
  owall* W = &e->wl[G->wall];
  const v3 F = V(Fa[0], Fa[1], Fa[2]);
  const v3 O_f = vdiv(F, vnorm(F));
  if (!G->bonded) {
    int it = 0;
    double d, df, ef;
    do {
      double k = 0.0;
      for (int c = 0; c < G->n; c++) {
        const opart* p = &e->P[G->pidx[c]];
        double dist = vdot(vsub(p->pos, W->origin), W->normal);
        if (dist < p->rad) k += G->keff[c];
      }
      if (k > 0) {
        v3 Fl = V(0.0, 0.0, 0.0);
        for (int c = 0; c < G->n; c++) {
          if (!G->inner[c]) continue;
          const opart* p = &e->P[G->pidx[c]];
          v3 f = V(0.0, 0.0, 0.0);
          double dist = vdot(vsub(p->pos, W->origin), W->normal);
          if (dist < p->rad) f = vmul(W->normal, G->keff[c] * (p->rad - dist));
          f = vmul(f, -1.0);
          Fl = vadd(Fl, vmul(O_f, vdot(f, O_f)));
        }
        df = vdot(vadd(F, Fl), O_f);
        d = df / k;
        ef = df / vnorm(F);
      } else {
        d = 1e-5;
        ef = 1;
      }
      W->origin = vadd(W->origin, vmul(O_f, d));
      it++;
    } while ((it < 50) && (ef > 1e-3));
    return ESP_OK;
  }
  double scale_base = 3.1415926536;
  double K = 0.0;
  for (int c = 0; c < G->n; c++) { /* bonded interactions first, then elastic */
    if (!G->tagged[c] || !G->inner[c]) continue;
    const opart* p = &e->P[G->pidx[c]];
    K += G->k * (e->do2d ? 1.0 : scale_base * p->rad);
  }
  for (int c = 0; c < G->n; c++) {
    if (G->tagged[c] || !G->inner[c]) continue;
    const opart* p = &e->P[G->pidx[c]];
    double dist = vdot(vsub(p->pos, W->origin), W->normal);
    if (dist < p->rad) K += G->keff[c];
  }
  if (!(K > 0)) return fail(e, ESP_ERR_NUMERIC, "wall group '%s' has no stiffness to servo", name);
  int it = 0;
  double d;
  do {
    v3 Fl = V(0.0, 0.0, 0.0);
    for (int c = 0; c < G->n; c++) {
      if (!G->tagged[c] || !G->inner[c]) continue;
      const opart* p = &e->P[G->pidx[c]];
      v3 D = vsub(vsub(p->pos, p->initpos), vsub(W->origin, W->oldpos));
      v3 f = vmul(D, G->k * (e->do2d ? 1.0 : scale_base * p->rad));
      Fl = vadd(Fl, vmul(O_f, vdot(f, O_f)));
    }
    for (int c = 0; c < G->n; c++) {
      if (G->tagged[c] || !G->inner[c]) continue;
      const opart* p = &e->P[G->pidx[c]];
      v3 f = V(0.0, 0.0, 0.0);
      double dist = vdot(vsub(p->pos, W->origin), W->normal);
      if (dist < p->rad) f = vmul(W->normal, G->keff[c] * (p->rad - dist));
      f = vmul(f, -1.0);
      Fl = vadd(Fl, vmul(O_f, vdot(f, O_f)));
    }
    d = vdot(vadd(F, Fl), O_f) / K;
    W->origin = vadd(W->origin, vmul(O_f, d));
    it++;
  } while ((it < 10) && (fabs(d) > 10e-6));
  return ESP_OK;
}

int eso_set_wall_step_motion(eso_engine* e, const char* name, const double d[3])
{
  int w = find_wall(e, name);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", name ? name : "");
  e->wl[w].step_motion = V(d[0], d[1], d[2]);
  e->wl[w].has_motion = (d[0] != 0.0 || d[1] != 0.0 || d[2] != 0.0);
  return ESP_OK;
}

int eso_run_time_steps(eso_engine* e, int nsteps)
{
  if (nsteps < 0) return fail(e, ESP_ERR_ARG, "negative step count");
  for (int s = 0; s < nsteps; s++) run_one_step(e);
  return ESP_OK;
}

int eso_get_num_particles(eso_engine* e) { return e->np; }

/* LsmMpiPy::getTimeStep (LsmMpiPy.cpp:466-468): the steps completed */
int eso_get_time_step(eso_engine* e, long long* out)
{
  if (!out) return ESP_ERR_ARG;
  *out = e->t;
  return ESP_OK;
}

int eso_get_stats(eso_engine* e, esp_stats* o)
{
  memset(o, 0, sizeof *o);
  o->time_step = e->t;
  o->num_rebuilds = e->timestamp;
  o->num_particles = e->np;
  for (int g = 0; g < e->nbg; g++) o->num_bonds += e->bg[g].n;
  for (int g = 0; g < e->ndg; g++) o->num_dynamic += e->dg[g].n;
  o->num_candidates = e->nlast;
  o->num_broken = e->nbrk / 3;
  o->num_host_phases = e->num_host_phases;
  return ESP_OK;
}

int eso_get_particle_ids(eso_engine* e, int* out)
{
  for (int k = 0; k < e->nidkey; k++) out[k] = (int)(e->idkey[k] >> 32);
  return ESP_OK;
}

int eso_get_particle_data(eso_engine* e, int field, double* out)
{
  for (int k = 0; k < e->nidkey; k++) {
    const opart* p = &e->P[e->idkey[k] & 0xffffffffLL];
    switch (field) {
      case ESP_FIELD_POS: out[3 * k] = p->pos.x; out[3 * k + 1] = p->pos.y; out[3 * k + 2] = p->pos.z; break;
      case ESP_FIELD_VEL: out[3 * k] = p->vel.x; out[3 * k + 1] = p->vel.y; out[3 * k + 2] = p->vel.z; break;
      case ESP_FIELD_ANGVEL: out[3 * k] = p->angvel.x; out[3 * k + 1] = p->angvel.y; out[3 * k + 2] = p->angvel.z; break;
      case ESP_FIELD_QUAT: out[4 * k] = p->q.s; out[4 * k + 1] = p->q.v.x; out[4 * k + 2] = p->q.v.y; out[4 * k + 3] = p->q.v.z; break;
      case ESP_FIELD_FORCE: out[3 * k] = p->force.x; out[3 * k + 1] = p->force.y; out[3 * k + 2] = p->force.z; break;
      case ESP_FIELD_MOMENT: out[3 * k] = p->moment.x; out[3 * k + 1] = p->moment.y; out[3 * k + 2] = p->moment.z; break;
      case ESP_FIELD_RAD: out[k] = p->rad; break;
      case ESP_FIELD_MASS: out[k] = p->mass; break;
      case ESP_FIELD_OLDPOS: out[3 * k] = p->oldpos.x; out[3 * k + 1] = p->oldpos.y; out[3 * k + 2] = p->oldpos.z; break;
      default: return fail(e, ESP_ERR_ARG, "unknown field %d", field);
    }
  }
  return ESP_OK;
}

/* Restart checkpoints are a file format, not part of the per-step path this
 * restatement checks; the engine's writer and reader are tested on their own. */
/* Test instrumentation: per particle (ascending id) the sum of the norms of
 * the forces and the sum of the norms of the moments applied in the last step,
 * sum_k |f_ik| and sum_k |m_ik| (the scale of a normwise per-particle
 * tolerance). */
int eso_get_force_scale(eso_engine* e, double* out)
{
  for (int k = 0; k < e->nidkey; k++) {
    const opart* p = &e->P[e->idkey[k] & 0xffffffffLL];
    out[2 * k] = p->fabs_sum;
    out[2 * k + 1] = p->mabs_sum;
  }
  return ESP_OK;
}

int eso_save_checkpoint(eso_engine* e, const char* prefix, int num_time_steps, int precision)
{
  (void)prefix; (void)num_time_steps; (void)precision;
  return fail(e, ESP_ERR_STATE, "checkpoints are not part of the oracle");
}

int eso_load_checkpoint(eso_engine* e, const char* info_file)
{
  (void)info_file;
  return fail(e, ESP_ERR_STATE, "checkpoints are not part of the oracle");
}

int eso_get_particle_stress(eso_engine* e, double* out)
{
  for (int k = 0; k < e->nidkey; k++) {
    const opart* p = &e->P[e->idkey[k] & 0xffffffffLL];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) out[9 * k + 3 * i + j] = p->sigma[i][j];
  }
  return ESP_OK;
}

int eso_get_wall_position(eso_engine* e, const char* name, double out[3])
{
  int w = find_wall(e, name);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", name ? name : "");
  out[0] = e->wl[w].origin.x; out[1] = e->wl[w].origin.y; out[2] = e->wl[w].origin.z;
  return ESP_OK;
}

int eso_get_wall_force(eso_engine* e, const char* name, double out[3])
{
  int w = find_wall(e, name);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", name ? name : "");
  out[0] = e->wl[w].force.x; out[1] = e->wl[w].force.y; out[2] = e->wl[w].force.z;
  return ESP_OK;
}

int eso_get_wall_force_history(eso_engine* e, const char* name, int n, double* out)
{
  int w = find_wall(e, name);
  if (w < 0) return fail(e, ESP_ERR_ARG, "unknown wall '%s'", name ? name : "");
  int have = e->wl[w].nhist / 3;
  int m = n < have ? n : have;
  if (out) memcpy(out, e->wl[w].hist + 3 * (have - m), sizeof(double) * 3 * (size_t)m);
  return m;
}

static int group_sorted(eso_engine* e, const char* name, long long** keys_out, int** perm_out)
{
  int g = find_dg(e, name);
  int n;
  long long* keys;
  if (g >= 0) {
    n = e->dg[g].n;
    keys = malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) keys[i] = pkey(e->dg[g].v[i].id1, e->dg[g].v[i].id2);
  } else {
    g = find_bg(e, name);
    if (g < 0) return -1;
    n = e->bg[g].n;
    keys = malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) keys[i] = pkey(e->bg[g].v[i].id1, e->bg[g].v[i].id2);
  }
  keyidx* a = malloc(sizeof(keyidx) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) { a[i].k = keys[i]; a[i].i = i; }
  sort_keyidx(a, n);
  int* perm = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) perm[i] = a[i].i;
  free(a);
  *keys_out = keys;
  *perm_out = perm;
  return n;
}

int eso_get_interaction_count(eso_engine* e, const char* name)
{
  int g = find_dg(e, name);
  if (g >= 0) return e->dg[g].n;
  g = find_bg(e, name);
  if (g >= 0) return e->bg[g].n;
  return -1;
}

int eso_get_interactions(eso_engine* e, const char* name, int* id1, int* id2)
{
  long long* keys;
  int* perm;
  int n = group_sorted(e, name, &keys, &perm);
  if (n < 0) return fail(e, ESP_ERR_ARG, "unknown interaction group '%s'", name ? name : "");
  for (int i = 0; i < n; i++) {
    id1[i] = (int)(keys[perm[i]] >> 32);
    id2[i] = (int)(keys[perm[i]] & 0xffffffffLL);
  }
  free(keys);
  free(perm);
  return n;
}

int eso_set_interaction_group_tags(eso_engine* e, const char* name, int tag1, int mask1, int tag2,
                                   int mask2)
{
  int g = find_dg(e, name);
  if (g < 0) return fail(e, ESP_ERR_ARG, "not a dynamic interaction group");
  ogroup* G = &e->dg[g];
  G->tagged = 1;
  if (tag1 <= tag2) {
    G->tag1 = tag1; G->mask1 = mask1; G->tag2 = tag2; G->mask2 = mask2;
  } else {
    G->tag1 = tag2; G->mask1 = mask2; G->tag2 = tag1; G->mask2 = mask1;
  }
  return ESP_OK;
}

int eso_track_bond_criteria(eso_engine* e)
{
  (void)e;
  return ESP_OK;
}

/* CRotBondedInteraction::getCriterion (RotBondedInteraction.cpp:221-227) on the
 * values the last calcForces cached, bonds in list order */
int eso_get_bond_criteria(eso_engine* e, const char* name, int* id1, int* id2, double* crit)
{
  int g = find_bg(e, name);
  if (g < 0) return fail(e, ESP_ERR_ARG, "not a bonded interaction group");
  const obgroup* G = &e->bg[g];
  if (id1 && id2 && crit)
    for (int i = 0; i < G->n; i++) {
      const obond* b = &G->v[i];
      id1[i] = b->id1;
      id2[i] = b->id2;
      crit[i] = b->nForce / b->max_n + b->shForce / b->max_s + b->tMoment / b->max_t +
                b->bMoment / b->max_b;
    }
  return G->n;
}

int eso_get_group_list_order(eso_engine* e, const char* name, int* id1, int* id2)
{
  int g = find_dg(e, name);
  if (g >= 0) {
    for (int i = 0; i < e->dg[g].n; i++) { id1[i] = e->dg[g].v[i].id1; id2[i] = e->dg[g].v[i].id2; }
    return e->dg[g].n;
  }
  g = find_bg(e, name);
  if (g >= 0) {
    for (int i = 0; i < e->bg[g].n; i++) { id1[i] = e->bg[g].v[i].id1; id2[i] = e->bg[g].v[i].id2; }
    return e->bg[g].n;
  }
  return fail(e, ESP_ERR_ARG, "unknown interaction group '%s'", name ? name : "");
}

int eso_get_friction_state(eso_engine* e, const char* name, double* ffric, int* flags)
{
  int g = find_dg(e, name);
  if (g < 0 || e->dg[g].kind != G_ED_ROTFRICTION)
    return fail(e, ESP_ERR_ARG, "'%s' is not a RotFriction group", name ? name : "");
  long long* keys;
  int* perm;
  int n = group_sorted(e, name, &keys, &perm);
  for (int i = 0; i < n; i++) {
    const odyn* d = &e->dg[g].v[perm[i]];
    ffric[3 * i] = d->ffric.x;
    ffric[3 * i + 1] = d->ffric.y;
    ffric[3 * i + 2] = d->ffric.z;
    flags[i] = (d->slipping ? 1 : 0) | (d->touching ? 2 : 0);
  }
  free(keys);
  free(perm);
  return n;
}

int eso_get_neighbour_pairs(eso_engine* e, int* id1, int* id2)
{
  if (id1 && id2)
    for (int i = 0; i < e->nlast; i++) {
      id1[i] = (int)(e->last_full[i] >> 32);
      id2[i] = (int)(e->last_full[i] & 0xffffffffLL);
    }
  return e->nlast;
}

int eso_get_bond_breaks(eso_engine* e, int* step, int* id1, int* id2)
{
  int n = e->nbrk / 3;
  if (step && id1 && id2) {
    /* events are logged in step order; within a step sort by (id1, id2) */
    typedef struct { int s, a, b; } ev;
    ev* v = malloc(sizeof(ev) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) { v[i].s = e->brk[3 * i]; v[i].a = e->brk[3 * i + 1]; v[i].b = e->brk[3 * i + 2]; }
    for (int i = 1; i < n; i++) {
      ev x = v[i];
      int j = i - 1;
      while (j >= 0 && (v[j].s > x.s || (v[j].s == x.s && (v[j].a > x.a || (v[j].a == x.a && v[j].b > x.b))))) { v[j + 1] = v[j]; j--; }
      v[j + 1] = x;
    }
    for (int i = 0; i < n; i++) { step[i] = v[i].s; id1[i] = v[i].a; id2[i] = v[i].b; }
    free(v);
  }
  return n;
}

int eso_track_pair_events(eso_engine* e, const char* name)
{
  int g = find_dg(e, name ? name : "");
  if (g < 0) return fail(e, ESP_ERR_ARG, "no dynamic interaction group '%s'", name ? name : "");
  if (!e->first_time) return fail(e, ESP_ERR_STATE, "pair events are tracked from the first time step on");
  e->dg[g].track = 1;
  return ESP_OK;
}

int eso_get_pair_events(eso_engine* e, const char* name, int* step, int* id1, int* id2, int* kind)
{
  int g = find_dg(e, name ? name : "");
  if (g < 0) return fail(e, ESP_ERR_ARG, "no dynamic interaction group '%s'", name ? name : "");
  ogroup* G = &e->dg[g];
  if (!G->track) return fail(e, ESP_ERR_STATE, "pair events of '%s' are not tracked", name);
  int n = G->nev / 4;
  /* logged in step order and, within a step, in (id1, id2) order */
  if (step && id1 && id2 && kind)
    for (int i = 0; i < n; i++) {
      step[i] = G->ev[4 * i];
      id1[i] = G->ev[4 * i + 1];
      id2[i] = G->ev[4 * i + 2];
      kind[i] = G->ev[4 * i + 3];
    }
  return n;
}

/* The oracle has no device kernels; timing reports zeros. */
int eso_set_kernel_timing(eso_engine* e, int enable)
{
  (void)e;
  (void)enable;
  return ESP_OK;
}

int eso_get_kernel_times(eso_engine* e, double* ms, long long* launches)
{
  (void)e;
  for (int k = 0; k < ESP_KCLASS_COUNT; k++) {
    if (ms) ms[k] = 0.0;
    if (launches) launches[k] = 0;
  }
  return ESP_OK;
}

int eso_get_rebuild_steps(eso_engine* e, int* steps)
{
  if (steps) memcpy(steps, e->rebuild_steps, sizeof(int) * (size_t)e->nreb);
  return e->nreb;
}

/* ------------------------------------------------------------------------- */
/* Single-law entry points                                                     */
/* ------------------------------------------------------------------------- */
void eso_law_rot_bond(const double* p, const double* b, double* out)
{
  v3 pos1 = V(p[0], p[1], p[2]);
  quat q1 = {p[3], {p[4], p[5], p[6]}};
  double r1 = p[7];
  v3 pos2 = V(p[8], p[9], p[10]);
  quat q2 = {p[11], {p[12], p[13], p[14]}};
  double r2 = p[15];
  bond_out o;
  rot_bond_law(pos1, q1, r1, pos2, q2, r2, b[0], b[1], b[2], b[3], V(b[4], b[5], b[6]), &o);
  double* w = out;
  *w++ = o.f1.x; *w++ = o.f1.y; *w++ = o.f1.z;
  *w++ = o.m1.x; *w++ = o.m1.y; *w++ = o.m1.z;
  *w++ = o.f2.x; *w++ = o.f2.y; *w++ = o.f2.z;
  *w++ = o.m2.x; *w++ = o.m2.y; *w++ = o.m2.z;
  *w++ = o.nForce; *w++ = o.shForce; *w++ = o.tMoment; *w++ = o.bMoment;
}

void eso_law_rot_friction(const double* p, const double* prm, double* state, double* out)
{
  v3 ff = V(state[0], state[1], state[2]);
  int slipping = state[3] != 0.0;
  double ediss = 0;
  v3 fn;
  fric_out o;
  rot_friction_law(V(p[0], p[1], p[2]), V(p[3], p[4], p[5]), V(p[6], p[7], p[8]), p[9],
                   V(p[10], p[11], p[12]), V(p[13], p[14], p[15]), V(p[16], p[17], p[18]), p[19],
                   prm[0], prm[1], prm[2], prm[3], prm[4], &ff, &slipping, &ediss, &fn, &o);
  state[0] = ff.x; state[1] = ff.y; state[2] = ff.z; state[3] = slipping;
  for (int i = 0; i < 13; i++) out[i] = 0.0;
  if (o.touching) {
    /* particle 1 receives -Fn then +Ffric; particle 2 receives +Fn then -Ffric */
    v3 f1 = vadd(vmul(fn, -1.0), o.f1);
    v3 f2 = vadd(fn, o.f2);
    out[0] = f1.x; out[1] = f1.y; out[2] = f1.z;
    out[3] = o.m1.x; out[4] = o.m1.y; out[5] = o.m1.z;
    out[6] = f2.x; out[7] = f2.y; out[8] = f2.z;
    out[9] = o.m2.x; out[10] = o.m2.y; out[11] = o.m2.z;
    out[12] = 1.0;
  }
}

/* The rigid form; state[4] is the touching flag of the last step. */
void eso_law_rot_friction_rigid(const double* p, const double* prm, double* state, double* out)
{
  v3 ff = V(state[0], state[1], state[2]);
  int slipping = state[3] != 0.0;
  double ediss = 0;
  v3 fn;
  fric_out o;
  rot_friction_rigid_law(V(p[0], p[1], p[2]), V(p[3], p[4], p[5]), V(p[6], p[7], p[8]), p[9],
                         V(p[10], p[11], p[12]), V(p[13], p[14], p[15]), V(p[16], p[17], p[18]),
                         p[19], prm[0], prm[1], prm[2], prm[3], prm[4], state[4] != 0.0, &ff,
                         &slipping, &ediss, &fn, &o);
  state[0] = ff.x; state[1] = ff.y; state[2] = ff.z; state[3] = slipping; state[4] = o.touching;
  for (int i = 0; i < 13; i++) out[i] = 0.0;
  if (o.touching) {
    v3 f1 = vadd(vmul(fn, -1.0), o.f1);
    v3 f2 = vadd(fn, o.f2);
    out[0] = f1.x; out[1] = f1.y; out[2] = f1.z;
    out[3] = o.m1.x; out[4] = o.m1.y; out[5] = o.m1.z;
    out[6] = f2.x; out[7] = f2.y; out[8] = f2.z;
    out[9] = o.m2.x; out[10] = o.m2.y; out[11] = o.m2.z;
    out[12] = 1.0;
  }
}

void eso_law_nrot_elastic(const double* p, double k, double* out)
{
  v3 p1 = V(p[0], p[1], p[2]), p2 = V(p[4], p[5], p[6]);
  double r1 = p[3], r2 = p[7];
  v3 D = vsub(p1, p2);
  double dist = vdot(D, D);
  double eq = r1 + r2;
  for (int i = 0; i < 7; i++) out[i] = 0.0;
  if (dist < eq * eq) {
    dist = sqrt(dist);
    v3 force = vmul(D, k * (dist - eq) / dist);
    out[0] = -force.x; out[1] = -force.y; out[2] = -force.z;
    out[3] = force.x; out[4] = force.y; out[5] = force.z;
    out[6] = 1.0;
  }
}

void eso_bond_constants(const double* in, const double* prm, double* out)
{
  esp_rot_bond_prms bp;
  memset(&bp, 0, sizeof bp);
  bp.kr = prm[0]; bp.ks = prm[1]; bp.kt = prm[2]; bp.kb = prm[3];
  bp.scaling = prm[8] != 0.0;
  bp.meanR_scaling = prm[9] != 0.0;
  obond b;
  bond_constants(V(in[0], in[1], in[2]), in[3], V(in[4], in[5], in[6]), in[7], &bp, prm + 4,
                 prm[10] != 0.0, &b);
  out[0] = b.kr; out[1] = b.ks; out[2] = b.kt; out[3] = b.kb;
  out[4] = b.max_n; out[5] = b.max_s; out[6] = b.max_t; out[7] = b.max_b;
  out[8] = b.D.x; out[9] = b.D.y; out[10] = b.D.z;
}

/* ------------------------------------------------------------------------- */
/* Math hooks: the oracle's own primitives, for comparison with the          */
/* reference's compiled Foundation classes (oracle/_ref/libref_math.so).      */
/* ------------------------------------------------------------------------- */
static quat q_in(const double* a) { quat q = {a[0], {a[1], a[2], a[3]}}; return q; }
static void q_out(quat q, double* o) { o[0] = q.s; o[1] = q.v.x; o[2] = q.v.y; o[3] = q.v.z; }
static mat3 m_in(const double* a)
{
  mat3 m;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m.m[i][j] = a[3 * i + j];
  return m;
}
static void m_out(const mat3* m, double* o)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[3 * i + j] = m->m[i][j];
}
void eso_math_quat_mul(const double* a, const double* b, double* o) { q_out(qmul(q_in(a), q_in(b)), o); }
void eso_math_quat_inverse(const double* a, double* o) { q_out(qinv(q_in(a)), o); }
void eso_math_quat_to_matrix(const double* a, double* o) { mat3 m = qmat(q_in(a)); m_out(&m, o); }
void eso_math_mat_vec(const double* m, const double* v, double* o)
{
  mat3 M = m_in(m);
  v3 r = mmul(&M, V(v[0], v[1], v[2]));
  o[0] = r.x; o[1] = r.y; o[2] = r.z;
}
void eso_math_mat_trans(const double* m, double* o) { mat3 M = m_in(m); mat3 T = mtrans(&M); m_out(&T, o); }
void eso_math_vec_cross(const double* a, const double* b, double* o)
{
  v3 r = vcross(V(a[0], a[1], a[2]), V(b[0], b[1], b[2]));
  o[0] = r.x; o[1] = r.y; o[2] = r.z;
}
double eso_math_vec_dot(const double* a, const double* b) { return vdot(V(a[0], a[1], a[2]), V(b[0], b[1], b[2])); }
double eso_math_vec_norm(const double* a) { return vnorm(V(a[0], a[1], a[2])); }
/* quaternion part of integrate_particle, without the rescale */
void eso_math_quat_step(const double* q, const double* w, double dt, double* o)
{
  quat Q = q_in(q);
  quat W = {0, V(w[0], w[1], w[2])};
  quat dq = qmul(W, Q);
  double h = dt / 2.0;
  Q.s += h * dq.s;
  Q.v = vadd(Q.v, vmul(dq.v, h));
  q_out(Q, o);
}
/* One particle step of integrate_particle for a single rotational particle:
 * state = {pos3, vel3, angvel3, quat4, force3, moment3, inv_mass, inv_inert}
 * (22 doubles) updated in place. */
void eso_math_rot_particle_step(double* st, double dt)
{
  opart p;
  memset(&p, 0, sizeof p);
  p.pos = V(st[0], st[1], st[2]);
  p.vel = V(st[3], st[4], st[5]);
  p.angvel = V(st[6], st[7], st[8]);
  p.q = q_in(st + 9);
  p.force = V(st[13], st[14], st[15]);
  p.moment = V(st[16], st[17], st[18]);
  p.inv_mass = st[19];
  p.inv_inert = st[20];
  p.is_dyn = 1;
  p.is_rot = 1;
  integrate_particle(&p, dt, 1, 0);
  st[0] = p.pos.x; st[1] = p.pos.y; st[2] = p.pos.z;
  st[3] = p.vel.x; st[4] = p.vel.y; st[5] = p.vel.z;
  st[6] = p.angvel.x; st[7] = p.angvel.y; st[8] = p.angvel.z;
  q_out(p.q, st + 9);
}
