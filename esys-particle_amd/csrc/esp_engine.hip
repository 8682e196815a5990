// esp_engine.hip - host side of the engine: the C-ABI of include/esp.h and the
// per-step orchestration of CLatticeMaster::runOneStep / TSubLattice
// (LatticeMaster.cpp:2602-2725, SubLattice.hpp:2224-2536) over the kernels.
//
// Stepping.  A step needs host bookkeeping only when the neighbour table must
// be rebuilt (a particle moved more than verletDist/2 since the last rebuild),
// when bonds broke in the previous step, or on the first step.  All other steps
// are pure device work, so the host queues `steps_per_batch` of them at once;
// each queued kernel checks the previous step's status on the device and stops
// the queue when bookkeeping is due (the halt is recorded, later steps return).
// After the batch the host synchronises once, performs the bookkeeping for the
// halted step and continues.  The decisions are the reference's, made at the
// same steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <fstream>
#include <mutex>
#include <sstream>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/esp.h"
#include "esp_internal.h"
#include "esp_math.h"

namespace {

struct EspError : std::runtime_error {
  int code;
  EspError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw EspError(ESP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
  DBuf& operator=(DBuf&& o) noexcept
  {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    return *this;
  }
  ~DBuf()
  {
    if (p) (void)hipFree(p);
  }
  // capacity for n elements; contents are not kept
  void reserve(size_t n)
  {
    if (n <= cap && p) return;
    if (p) {
      HIPCHK(hipDeviceSynchronize());  // queued kernels may still read the old buffer
      HIPCHK(hipFree(p));
    }
    p = nullptr;
    size_t c = std::max<size_t>(n + n / 4, 64);
    HIPCHK(hipMalloc(&p, c * sizeof(T)));
    cap = c;
  }
  void swap(DBuf& o)
  {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

// page-locked host memory (hipHostMalloc): device-to-host copies into it run
// at full link speed instead of through the runtime's staging buffers
template <class T>
struct PinBuf {
  T* p = nullptr;
  size_t cap = 0;
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
  ~PinBuf()
  {
    if (p) (void)hipHostFree(p);
  }
  // capacity for n elements; contents are not kept
  void reserve(size_t n)
  {
    if (n <= cap && p) return;
    if (p) {
      HIPCHK(hipDeviceSynchronize());  // a queued copy may still write the old buffer
      HIPCHK(hipHostFree(p));
    }
    p = nullptr;
    size_t c = std::max<size_t>(n + n / 4, 64);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p), c * sizeof(T), hipHostMallocDefault));
    cap = c;
  }
};

struct DynGroup {
  std::string name;
  int kind;  // 1 NRotElastic, 2 RotElastic, 3 RotFriction
  double k = 0, ks = 0, mu_d = 0, mu_s = 0;
  int rigid = 0;  // RotFriction rigid=True
  int scaling = 1, meanR = 1;
  int n = 0;  // main list sorted by slot key; the appendix (device count napp) follows it
  DBuf<int2> ij;
  DBuf<long long> key;
  DBuf<double4> ff;   // shear histories the next step reads
  DBuf<double4> ff2;  // the other history buffer (the next step writes it)
  DBuf<int> napp;     // [0] pairs appended, [1] pairs the force blocks cover
  DBuf<unsigned long long> hs;  // stored pairs (history groups)
  unsigned hs_mask = 0;
  long long update_ts = 0;
  std::vector<int> ex;
  int tagged = 0, tag1 = 0, mask1 = 0, tag2 = 0, mask2 = 0;  // createInteractionGroupTagged
  int rec_off = 0, app_rec_off = 0;
  // pairs and shear histories loaded from a restart checkpoint, stored at the
  // first step (PIS_ED::loadCheckPointData, pi_storage_ed.hpp:225-255)
  std::vector<int> pre_id1, pre_id2;
  std::vector<double4> pre_ff;
  // pair-event log (esp_track_pair_events): the stored pairs as sorted id keys
  // at the last logged update, appendix pairs logged since then (not yet in
  // ev_prev), the appendix entries already logged, each appendix entry's step
  // of creation, and the events (step, id1, id2, kind)
  int track = 0;
  std::vector<long long> ev_prev, ev_pend;
  int app_logged = 0;
  DBuf<int> app_step;
  std::vector<int4> ev;
};

struct BondGroup {
  std::string name;
  esp_rot_bond_prms prm;
  int n = 0;
  DBuf<int2> ij;  // dead bonds hold (-1, -1) until the next compaction
  DBuf<double4> k4, mx4, D4;
  DBuf<int> rank;  // position in connection order (the order bonded forces are summed in)
  DBuf<unsigned long long> hs;  // keys of live bonds (exclusion lookups)
  unsigned hs_mask = 0;
  std::vector<int> hid1, hid2;  // ids in storage order (sorted by particle slots)
  std::vector<double4> h_k4, h_mx4, h_D4;  // staged until the first step
  int rec_off = 0, bond_off = 0;
  // BrittleBeamSC: broken by the particle stresses (tension cutoff, cohesion,
  // tan of the friction angle) instead of the bond forces
  int sc = 0;
  double sc_sigma3 = 0, sc_cohesion = 0, sc_tan = 0;
};

struct Wall {
  std::string name;
  double o[3], n[3], step[3];
  double old[3];  // CWall::m_oldpos (Wall.cpp:38): the origin at creation
  bool has_motion = false;
  std::vector<double> hist;
  double force[3] = {0, 0, 0};
};

struct WallGroup {
  std::string name;
  int wall;
  double k;
  int bonded = 0, tag = 0, mask = 0;  // CBWallInteractionGroup (tagged particles bonded)
  int tagged = 0;  // CTaggedEWallInteractionGroup (only tagged particles interact)
};

struct Gravity {
  std::string name;
  double g[3];
};

// a damping group; key: its place in the reference's damping map (the type
// name for Damping / RotDamping, the group name for the local forms)
struct Damping {
  int type;
  std::string key;
  double visc;
  int maxit;
};

inline long long pkey(int a, int b)
{
  return (long long)(((unsigned long long)(unsigned)a << 32) | (unsigned long long)(unsigned)b);
}

thread_local std::string g_create_err;

void* scratch_on(int dev);  // the allocation-free servo's staging (defined below)

}  // namespace

struct esp_engine {
  esp_config cfg;
  int rot = 0, do2d = 0;
  double dt = 0, alpha = 0;
  GridParams grid;
  int ncell = 0;
  hipStream_t stream = nullptr;
  int K = 32;
  int kcur = 1;

  // host staging (until the first step) and host mirrors
  std::vector<int> h_id, h_tag;
  std::vector<double> h_pos, h_rad, h_mass, h_vel;
  std::vector<unsigned> h_flags;
  // state a restart checkpoint may carry beyond createParticles (per slot)
  std::vector<double> h_quat, h_ang, h_inert;  // h_inert 0: 0.4 m r^2 (0.5 m r^2 in 2-D)
  std::vector<double> h_init;                  // CParticle::m_initpos
  // creation rank per slot: the particles' order in the reference's particle
  // list (insertion order; erasure keeps the order of the rest)
  std::vector<long long> h_order;
  long long next_order = 0;
  // applyForceToWall's candidates in particle-list order: the slot of every
  // creation order (-1: erased), rebuilt after the slots change, and the
  // staging buffers the candidates are packed into in that order
  DBuf<int> ord_slot;
  bool ord_valid = false;
  DBuf<int> wc_mark, wc_off;
  DBuf<WallCand> wc_cand;
  PinBuf<WallCand> hw_cand;
  bool forced_search = false;  // a checkpoint forced the neighbour search after the last step
  std::vector<int> h_conn;
  std::unordered_map<int, int> id2slot;
  bool started = false;
  int n = 0;

  DBuf<double4> pos, vel, ang, quat, old, frc, mom;
  DBuf<unsigned> flags;
  DBuf<int> id, cell;
  DBuf<int> cell_count, cell_start, cell_fill, cell_items;
  DBuf<int> cand_cnt, cand_off, cand_p, cand_q;
  DBuf<long long> cand_key;
  int ncand = 0;
  DBuf<int> keep, keep_off, src, scan_tmp;
  DBuf<int> np_p, np_q;  // flagged (new) pair list
  DBuf<long long> np_key;
  int nnew = 0;
  DBuf<int2> tmp_ij, tmp_ij2;
  DBuf<long long> tmp_key, tmp_key2;
  DBuf<double4> tmp_ff, tmp_a, tmp_b, tmp_c;
  DBuf<int> inc_cnt, inc_off, inc_fill, inc_ent, tmp_rank;
  DBuf<unsigned long long> inc_ent64;
  DBuf<int> xcnt, flist;
  DBuf<int4> evt;
  DBuf<UpdState> upd;
  int nrec = 0;
  DBuf<double> slots;     // per-particle slots: CSR runs, then ESP_XSLOTS appendix slots each
  DBuf<unsigned char> sact;  // per slot: 1 while it holds a nonzero share
  DBuf<int2> slot_of;     // record -> its two slots
  int ss = 6;             // doubles per slot (9 when particle stresses are tracked)
  int xbase = 0;          // first appendix slot
  DBuf<double> sig;       // stress sums [particle][9] when tracked
  DBuf<StepStatus> ring;
  int ring_size = 0;
  DBuf<int> halted, broken_list, dev_int;
  int broken_cap = 0;
  DBuf<double> wall_part, wall_hist;

  std::vector<DynGroup> dg;
  std::vector<BondGroup> bg;
  std::vector<Wall> walls;
  std::vector<WallGroup> wg;
  std::vector<Gravity> grav;
  std::vector<Damping> damp;  // sorted by key
  bool track_crit = false;    // bonds store their breaking criterion each step
  DBuf<double> crit;          // per global bond index
  DBuf<double> inert;         // moments of inertia, uploaded when a rotational damping exists
  DBuf<double4> init;         // positions at creation, uploaded when a bonded wall exists
  DBuf<int> tag;              // particle tags, uploaded when a tagged group exists
  std::vector<int> dord, bord, word, gord;

  long long t = 0, timestamp = 0, last_ns = -1;
  bool first_time = true, pending_host = false;
  int pending_reason = 0;
  long long num_host_phases = 0;
  std::vector<int> rebuild_steps;
  std::vector<int> breaks;  // step, id1, id2
  long long dead_pending = 0;  // bonds marked dead on the device, not yet compacted
  std::string err;

  // kernel timing: one (start, stop, class) record per launch of a queued step
  bool timing = false;
  struct TimedLaunch {
    hipEvent_t a, b;
    int cls;
    int step_id;
  };
  std::vector<TimedLaunch> tl;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double k_ms[ESP_KCLASS_COUNT] = {};
  long long k_n[ESP_KCLASS_COUNT] = {};

  PView view()
  {
    PView v;
    v.pos = pos.p; v.vel = vel.p; v.ang = ang.p; v.quat = quat.p; v.old = old.p;
    v.frc = frc.p; v.mom = mom.p; v.flags = flags.p; v.id = id.p; v.n = n;
    return v;
  }
};

namespace {

void sync(esp_engine* e) { HIPCHK(hipStreamSynchronize(e->stream)); }

hipEvent_t next_event(esp_engine* e)
{
  if (e->ev_used == e->ev_pool.size()) {
    hipEvent_t ev;
    HIPCHK(hipEventCreate(&ev));
    e->ev_pool.push_back(ev);
  }
  return e->ev_pool[e->ev_used++];
}

// (start, stop) events for one launch of a queued step when timing is on,
// else two nulls: the launch stamps them at the kernel's own start and end
struct LaunchEvents {
  hipEvent_t a = nullptr, b = nullptr;
};
LaunchEvents launch_events(esp_engine* e)
{
  LaunchEvents t;
  if (!e->timing) return t;
  t.a = next_event(e);
  t.b = next_event(e);
  return t;
}
// keeps the record of a launch that happened (a skipped launch stamps nothing)
void note_launch(esp_engine* e, const LaunchEvents& t, bool launched, int cls, int step_id)
{
  if (t.a && launched) e->tl.push_back({t.a, t.b, cls, step_id});
}

// after a synchronisation: add the launches of steps 1..done of the batch
void collect_timing(esp_engine* e, int done)
{
  if (!e->timing) return;
  for (auto& t : e->tl) {
    if (t.step_id > done) continue;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
    e->k_ms[t.cls] += ms;
    e->k_n[t.cls]++;
  }
  e->tl.clear();
  e->ev_used = 0;
}

template <class T>
void h2d(esp_engine* e, T* d, const T* h, size_t cnt)
{
  if (cnt) HIPCHK(hipMemcpyAsync(d, h, cnt * sizeof(T), hipMemcpyHostToDevice, e->stream));
}
template <class T>
void d2h(esp_engine* e, T* h, const T* d, size_t cnt)
{
  if (cnt) HIPCHK(hipMemcpyAsync(h, d, cnt * sizeof(T), hipMemcpyDeviceToHost, e->stream));
}
template <class T>
void zero(esp_engine* e, T* d, size_t cnt)
{
  if (cnt) HIPCHK(hipMemsetAsync(d, 0, cnt * sizeof(T), e->stream));
}

// exclusive scan of `cnt` ints, returns the total (synchronises)
int scan_total(esp_engine* e, const int* in, int* out, int cnt)
{
  e->scan_tmp.reserve(scan_tmp_size(cnt));
  launch_exclusive_scan(e->stream, in, out, cnt, e->scan_tmp.p);
  HIPCHK(hipGetLastError());
  int total = 0;
  d2h(e, &total, out + cnt, 1);
  sync(e);
  return total;
}

std::vector<int> name_order(size_t n, const std::function<const std::string&(int)>& nm)
{
  std::vector<int> o(n);
  for (size_t i = 0; i < n; i++) o[i] = (int)i;
  std::sort(o.begin(), o.end(), [&](int a, int b) { return nm(a) < nm(b); });
  return o;
}

void refresh_orders(esp_engine* e)
{
  e->dord = name_order(e->dg.size(), [&](int i) -> const std::string& { return e->dg[i].name; });
  e->bord = name_order(e->bg.size(), [&](int i) -> const std::string& { return e->bg[i].name; });
  e->word = name_order(e->wg.size(), [&](int i) -> const std::string& { return e->wg[i].name; });
  e->gord = name_order(e->grav.size(), [&](int i) -> const std::string& { return e->grav[i].name; });
}

bool name_taken(esp_engine* e, const std::string& n)
{
  for (auto& g : e->dg) if (g.name == n) return true;
  for (auto& g : e->bg) if (g.name == n) return true;
  for (auto& g : e->wg) if (g.name == n) return true;
  for (auto& g : e->grav) if (g.name == n) return true;
  return false;
}

int find_dg(esp_engine* e, const char* n)
{
  for (size_t i = 0; i < e->dg.size(); i++) if (e->dg[i].name == n) return (int)i;
  return -1;
}
int find_bg(esp_engine* e, const char* n)
{
  for (size_t i = 0; i < e->bg.size(); i++) if (e->bg[i].name == n) return (int)i;
  return -1;
}
int find_wall(esp_engine* e, const char* n)
{
  for (size_t i = 0; i < e->walls.size(); i++) if (e->walls[i].name == n) return (int)i;
  return -1;
}

void require_setup(esp_engine* e, const char* what)
{
  if (e->started)
    throw EspError(ESP_ERR_STATE, std::string(what) +
                                      " must be called before the first time step on this engine");
}

// NeighborTable::index on the host (same arithmetic as the device kernel)
int host_cell(esp_engine* e, double x, double y, double z)
{
  const GridParams& g = e->grid;
  double f[3] = {std::floor((x - g.x0) / g.dim), std::floor((y - g.y0) / g.dim),
                 std::floor((z - g.z0) / g.dim)};
  for (double v : f)
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return -1;
  int ix = (int)f[0] - g.gx, iy = (int)f[1] - g.gy, iz = (int)f[2] - g.gz;
  if (g.circ_x && x >= g.x0 && x < g.x0 + g.period) ix = ix < 0 ? 0 : (ix >= g.nx ? g.nx - 1 : ix);
  if (ix >= 0 && ix < g.nx && iy >= 0 && iy < g.ny && iz >= 0 && iz < g.nz)
    return g.ny * g.nz * ix + g.nz * iy + iz;
  return -1;
}

ExclSet excl_set(esp_engine* e, const DynGroup& G)
{
  ExclSet s;
  memset(&s, 0, sizeof s);
  for (int b : G.ex) {
    s.h[s.count].t = e->bg[b].hs.p;
    s.h[s.count].mask = e->bg[b].hs_mask;
    s.count++;
  }
  s.tagged = G.tagged;
  s.tag1 = G.tag1;
  s.mask1 = G.mask1;
  s.tag2 = G.tag2;
  s.mask2 = G.mask2;
  s.tag = e->tag.p;
  return s;
}

unsigned hash_capacity(size_t n)
{
  size_t c = 1024;
  while (c < 2 * n + 2) c <<= 1;
  return (unsigned)c;
}

// (re)build a hash set over `cnt` keys (device array)
void build_hash(esp_engine* e, DBuf<unsigned long long>& hs, unsigned& mask, size_t capacity_for,
                const long long* keys, int cnt)
{
  const unsigned cap = hash_capacity(capacity_for);
  hs.reserve(cap);
  mask = cap - 1;
  HashRef h{hs.p, mask};
  launch_hash_clear(e->stream, h);
  launch_hash_insert(e->stream, h, keys, cnt);
  HIPCHK(hipGetLastError());
}

void rebuild_bond_hash(esp_engine* e, BondGroup& G)
{
  e->tmp_key.reserve(std::max(G.n, 1));
  launch_bond_keys(e->stream, G.n, G.ij.p, e->id.p, e->tmp_key.p);
  build_hash(e, G.hs, G.hs_mask, G.n, e->tmp_key.p, G.n);
}

// dynamic-group arrays hold the main list plus room for the appendix
void reserve_group(DynGroup& G, size_t n_main, bool hist)
{
  const size_t c = n_main + ESP_APP_CAP;
  G.ij.reserve(c);
  G.key.reserve(c);
  if (hist) G.ff.reserve(c);
}

void rebuild_presence(esp_engine* e, DynGroup& G)
{
  if (G.kind != 3) return;
  build_hash(e, G.hs, G.hs_mask, (size_t)G.n + ESP_APP_CAP, G.key.p, G.n);
}

// ---------------------------------------------------------------------------
// device upload at the first step
// ---------------------------------------------------------------------------
unsigned long long morton3(unsigned x, unsigned y, unsigned z)
{
  auto spread = [](unsigned long long v) {
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  return spread(x) | spread(y) << 1 | spread(z) << 2;
}

// Particles are stored in Morton order of their initial grid cells, so that
// interacting particles sit close together in memory.  The storage order
// changes no result: lists are keyed by ids or slots, and per-particle sums
// follow the reference's group and connection orders.
void spatial_order(esp_engine* e)
{
  const int n = e->n;
  const double o[3] = {e->grid.x0, e->grid.y0, e->grid.z0};
  std::vector<unsigned long long> code(n);
  for (int i = 0; i < n; i++) {
    unsigned c[3];
    for (int a = 0; a < 3; a++) {
      const double f = std::floor((e->h_pos[3 * i + a] - o[a]) / e->grid.dim) + 1.0;
      c[a] = (unsigned)std::max(0.0, std::min(f, 2097151.0));
    }
    code[i] = morton3(c[0], c[1], c[2]);
  }
  std::vector<int> perm(n);  // new slot -> old slot
  for (int i = 0; i < n; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return code[a] < code[b]; });
  auto permute = [&](auto& v, int w) {
    auto old = v;
    for (int i = 0; i < n; i++)
      for (int c = 0; c < w; c++) v[(size_t)w * i + c] = old[(size_t)w * perm[i] + c];
  };
  permute(e->h_id, 1);
  permute(e->h_tag, 1);
  permute(e->h_pos, 3);
  permute(e->h_vel, 3);
  permute(e->h_rad, 1);
  permute(e->h_mass, 1);
  permute(e->h_flags, 1);
  permute(e->h_init, 3);
  permute(e->h_quat, 4);
  permute(e->h_ang, 3);
  permute(e->h_inert, 1);
  permute(e->h_order, 1);
  e->ord_valid = false;
  e->id2slot.clear();
  for (int i = 0; i < n; i++) e->id2slot[e->h_id[i]] = i;
}

void upload(esp_engine* e)
{
  spatial_order(e);
  for (auto& G : e->bg)
    if (G.sc) e->ss = 9;
  const int n = e->n;
  const size_t cap = std::max(n, 1);
  e->pos.reserve(cap); e->vel.reserve(cap); e->old.reserve(cap); e->frc.reserve(cap);
  e->flags.reserve(cap); e->id.reserve(cap); e->cell.reserve(cap);
  if (e->rot) { e->ang.reserve(cap); e->quat.reserve(cap); e->mom.reserve(cap); }
  std::vector<double4> P(n), Vv(n), O(n), A(n), Q(n), Z(n);
  for (int i = 0; i < n; i++) {
    const double m = e->h_mass[i], r = e->h_rad[i];
    P[i] = make_double4(e->h_pos[3 * i], e->h_pos[3 * i + 1], e->h_pos[3 * i + 2], r);
    Vv[i] = make_double4(e->h_vel[3 * i], e->h_vel[3 * i + 1], e->h_vel[3 * i + 2],
                         (m != 0.0) ? 1.0 / m : 0.0);
    O[i] = make_double4(P[i].x, P[i].y, P[i].z, m);
    Z[i] = make_double4(0, 0, 0, 0);
    if (e->rot) {
      // CRotParticle: I = 0.4 m r^2 (3-D) / 0.5 m r^2 (2-D) (RotParticle.cpp:80-124, 160-177)
      const double I = e->h_inert[i] != 0.0 ? e->h_inert[i]
                                            : (e->do2d ? 0.5 * m * r * r : 0.4 * m * r * r);
      A[i] = make_double4(e->h_ang[3 * i], e->h_ang[3 * i + 1], e->h_ang[3 * i + 2], 1.0 / I);
      Q[i] = make_double4(e->h_quat[4 * i], e->h_quat[4 * i + 1], e->h_quat[4 * i + 2],
                          e->h_quat[4 * i + 3]);
    }
  }
  h2d(e, e->pos.p, P.data(), n);
  h2d(e, e->vel.p, Vv.data(), n);
  h2d(e, e->old.p, O.data(), n);
  h2d(e, e->frc.p, Z.data(), n);
  if (e->rot) {
    h2d(e, e->ang.p, A.data(), n);
    h2d(e, e->quat.p, Q.data(), n);
    h2d(e, e->mom.p, Z.data(), n);
  }
  // tags matching a bonded wall group (static: tags do not change)
  bool bonded_wall = false;
  for (size_t k = 0; k < e->word.size(); k++) {
    const WallGroup& G = e->wg[e->word[k]];
    if (!G.bonded && !G.tagged) continue;
    bonded_wall |= G.bonded != 0;
    for (int i = 0; i < n; i++)
      if ((e->h_tag[i] & G.mask) == (G.tag & G.mask)) e->h_flags[i] |= 1u << (PF_BTAG_SHIFT + k);
  }
  if (bonded_wall) {
    std::vector<double4> I0(n);
    for (int i = 0; i < n; i++)
      I0[i] = make_double4(e->h_init[3 * i], e->h_init[3 * i + 1], e->h_init[3 * i + 2], 0.0);
    e->init.reserve(cap);
    h2d(e, e->init.p, I0.data(), n);
    sync(e);  // the copy reads I0, which ends with this block
  }
  h2d(e, e->flags.p, e->h_flags.data(), n);
  h2d(e, e->id.p, e->h_id.data(), n);
  bool tagged_group = false;
  for (auto& G : e->dg) tagged_group |= G.tagged != 0;
  if (tagged_group) {
    e->tag.reserve(cap);
    h2d(e, e->tag.p, e->h_tag.data(), n);
  }
  bool rot_damping = false;
  for (auto& D : e->damp) rot_damping |= D.type == ESP_ROT_DAMPING;
  if (rot_damping) {
    std::vector<double> I(n);
    for (int i = 0; i < n; i++) {
      const double m = e->h_mass[i], r = e->h_rad[i];
      I[i] = e->h_inert[i] != 0.0 ? e->h_inert[i] : (e->do2d ? 0.5 * m * r * r : 0.4 * m * r * r);
    }
    e->inert.reserve(cap);
    h2d(e, e->inert.p, I.data(), n);
    sync(e);  // the copy reads I, which ends with this block
  }

  int total_bonds = 0;
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    G.bond_off = total_bonds;
    total_bonds += G.n;
    const size_t c = std::max(G.n, 1);
    G.ij.reserve(c); G.k4.reserve(c); G.mx4.reserve(c); G.D4.reserve(c); G.rank.reserve(c);
    if (G.n >= (1 << 27)) throw EspError(ESP_ERR_CAPACITY, "more than 2^27 bonds in one group");
    // stored in slot order; rank keeps the connection order for the force sums
    std::vector<int2> ij0(G.n);
    for (int b = 0; b < G.n; b++)
      ij0[b] = make_int2(e->id2slot[G.hid1[b]], e->id2slot[G.hid2[b]]);
    std::vector<int> rank(G.n);
    for (int b = 0; b < G.n; b++) rank[b] = b;
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) {
      return ij0[a].x != ij0[b].x ? ij0[a].x < ij0[b].x : ij0[a].y < ij0[b].y;
    });
    std::vector<int2> ij(G.n);
    std::vector<int> h1(G.n), h2(G.n);
    std::vector<double4> k4(G.n), mx4(G.n), D4(G.n);
    for (int b = 0; b < G.n; b++) {
      const int r = rank[b];
      ij[b] = ij0[r];
      h1[b] = G.hid1[r];
      h2[b] = G.hid2[r];
      k4[b] = G.h_k4[r];
      mx4[b] = G.h_mx4[r];
      D4[b] = G.h_D4[r];
    }
    G.hid1.swap(h1);
    G.hid2.swap(h2);
    h2d(e, G.ij.p, ij.data(), G.n);
    h2d(e, G.k4.p, k4.data(), G.n);
    h2d(e, G.mx4.p, mx4.data(), G.n);
    h2d(e, G.D4.p, D4.data(), G.n);
    h2d(e, G.rank.p, rank.data(), G.n);
    rebuild_bond_hash(e, G);
    sync(e);
    std::vector<double4>().swap(G.h_k4);
    std::vector<double4>().swap(G.h_mx4);
    std::vector<double4>().swap(G.h_D4);
  }
  for (auto& G : e->dg) {
    G.napp.reserve(2);
    zero(e, G.napp.p, 2);
    // pairs loaded from a checkpoint form the stored list the first update
    // starts from (their histories persist while touching)
    std::vector<std::pair<long long, int>> pk;
    for (size_t k = 0; k < G.pre_id1.size(); k++) {
      auto a = e->id2slot.find(G.pre_id1[k]), b = e->id2slot.find(G.pre_id2[k]);
      if (a == e->id2slot.end() || b == e->id2slot.end()) continue;
      pk.push_back({pkey(a->second, b->second), (int)k});
    }
    std::sort(pk.begin(), pk.end());
    pk.erase(std::unique(pk.begin(), pk.end(),
                         [](const std::pair<long long, int>& x, const std::pair<long long, int>& y) {
                           return x.first == y.first;
                         }),
             pk.end());
    const int np = (int)pk.size();
    reserve_group(G, np, G.kind == 3);
    if (np > 0) {
      std::vector<int2> ij(np);
      std::vector<long long> key(np);
      std::vector<double4> ff(np);
      for (int k = 0; k < np; k++) {
        key[k] = pk[k].first;
        ij[k] = make_int2((int)(pk[k].first >> 32), (int)(pk[k].first & 0xffffffffLL));
        ff[k] = G.pre_ff[pk[k].second];
      }
      h2d(e, G.ij.p, ij.data(), np);
      h2d(e, G.key.p, key.data(), np);
      if (G.kind == 3) h2d(e, G.ff.p, ff.data(), np);
      sync(e);  // the copies read ij, key and ff, which end with this block
    }
    G.n = np;
    rebuild_presence(e, G);
    if (G.track) {
      G.app_step.reserve(ESP_APP_CAP);
      G.ev_prev.clear();
      for (int k = 0; k < np; k++) {
        const int s = pk[k].second;
        G.ev_prev.push_back(pkey(G.pre_id1[s], G.pre_id2[s]));  // as dyn_view reads it
      }
      std::sort(G.ev_prev.begin(), G.ev_prev.end());
    }
  }
  e->broken_cap = std::max(total_bonds, 1);
  if (e->track_crit) {
    e->crit.reserve((size_t)e->broken_cap);
    zero(e, e->crit.p, (size_t)e->broken_cap);
  }
  e->broken_list.reserve(2 * (size_t)e->broken_cap);  // one list per step parity
  e->evt.reserve(e->broken_cap);
  e->ring_size = e->K + 2;
  e->ring.reserve(e->ring_size);
  zero(e, e->ring.p, e->ring_size);
  e->halted.reserve(1);
  zero(e, e->halted.p, 1);
  e->upd.reserve(1);
  zero(e, e->upd.p, 1);
  e->dev_int.reserve(4);
  e->cell_count.reserve(e->ncell + 1);
  e->cell_start.reserve(e->ncell + 1);
  e->cell_fill.reserve(e->ncell + 1);
  e->cell_items.reserve(cap);
  e->cand_cnt.reserve(cap + 1);
  e->cand_off.reserve(cap + 1);
  e->inc_cnt.reserve(cap + 1);
  e->inc_off.reserve(cap + 1);
  e->inc_fill.reserve(cap + 1);
  e->xcnt.reserve(cap);
  e->flist.reserve(cap);
  zero(e, e->xcnt.p, cap);
  const int nb = std::max(gather_blocks(n), 1);
  const int nwg = (int)e->wg.size();
  e->wall_part.reserve((size_t)(e->K + 1) * nb * std::max(nwg, 1) * 3);
  if (e->ss == 9) {
    e->sig.reserve((size_t)std::max(n, 1) * 9);
    zero(e, e->sig.p, (size_t)std::max(n, 1) * 9);
  }
  e->wall_hist.reserve((size_t)(e->K + 1) * std::max(nwg, 1) * 3);
  sync(e);
  e->started = true;
}

// ---------------------------------------------------------------------------
// neighbour search: ParallelParticleArray::rebuild + rebuildInteractions +
// resetDisplacements (pp_array.hpp:433-554, SubLattice.hpp:2360-2410)
// ---------------------------------------------------------------------------
// NeighborTable::build erases the particles outside the padded grid
// (ntable.hpp:281-286) and the interaction storages then drop every
// interaction that lost a particle (PIS_E::rebuild, pi_storage_e.hpp:111-155;
// PIS_NE / PIS_ED rebuild their lists from the new full list).  On the device:
// a stable compaction of every per-particle array in slot order, the pairs of
// every group remapped to the new slots (slots are renumbered monotonically, so
// key-sorted lists stay sorted and bonds keep their summation ranks), the hash
// sets rebuilt, the flagged-particle list remapped; the host mirrors follow
// with one copy of the keep flags.  Called on the host bookkeeping path after
// fold() (no dead bonds, no appendices), with e->cell holding -1 for the
// particles to erase.
int* broken_list_of(esp_engine* e, long long step);

template <class T>
void compact_buf(esp_engine* e, DBuf<T>& b, int n, int nk, int width)
{
  if (!b.p) return;
  DBuf<T> t;
  t.reserve((size_t)std::max(nk, 1) * width);
  launch_compact_field(e->stream, n, width, e->keep.p, e->keep_off.p, b.p, t.p);
  HIPCHK(hipGetLastError());
  sync(e);
  b.swap(t);
}

template <class T>
void compact_host(std::vector<T>& v, const std::vector<int>& keep, int width)
{
  size_t w = 0;
  for (size_t i = 0; i < keep.size(); i++)
    if (keep[i])
      for (int c = 0; c < width; c++) v[w++] = v[(size_t)width * i + c];
  v.resize(w);
}

void erase_outside(esp_engine* e)
{
  const int n = e->n;
  e->keep.reserve((size_t)n + 1);
  e->keep_off.reserve((size_t)n + 1);
  launch_keep_cells(e->stream, n, e->cell.p, e->keep.p);
  const int nk = scan_total(e, e->keep.p, e->keep_off.p, n);
  std::vector<int> hkeep(n), hslot(n);
  d2h(e, hkeep.data(), e->keep.p, n);
  DBuf<int> nslot;
  nslot.reserve((size_t)std::max(n, 1));
  launch_new_slot(e->stream, n, e->keep.p, e->keep_off.p, nslot.p);
  d2h(e, hslot.data(), nslot.p, n);
  sync(e);
  // Bonds whose break criterion held in the last step: the next step's update
  // breaks them by global bond index (bond_off + index in the group), so the
  // list is re-indexed below; bonds the erasure drops leave it (no break event
  // for an interaction that no longer exists).
  std::vector<int> pend;
  const size_t ring_i = (size_t)(e->t % e->ring_size);
  if (!e->first_time && e->t >= 1 && !e->bg.empty()) {
    StepStatus pst;
    d2h(e, &pst, e->ring.p + ring_i, 1);
    sync(e);
    if (pst.nbroken > 0) {
      pend.resize(std::min(pst.nbroken, e->broken_cap));
      d2h(e, pend.data(), broken_list_of(e, e->t), pend.size());
      sync(e);
    }
  }
  std::vector<int> old_off(e->bg.size()), old_n(e->bg.size());
  std::vector<std::vector<int>> bkeep(e->bg.size());
  for (size_t g = 0; g < e->bg.size(); g++) {
    old_off[g] = e->bg[g].bond_off;
    old_n[g] = e->bg[g].n;
  }
  // interactions first (they read the old slots through nslot)
  DBuf<int> pkeep, poff;
  DBuf<int2> rij;
  for (auto& G : e->bg) {
    if (G.n <= 0) continue;
    pkeep.reserve((size_t)G.n + 1);
    poff.reserve((size_t)G.n + 1);
    rij.reserve(G.n);
    launch_remap_pairs(e->stream, G.n, G.ij.p, nslot.p, rij.p, pkeep.p);
    const int nb2 = scan_total(e, pkeep.p, poff.p, G.n);
    std::vector<int>& bk = bkeep[&G - e->bg.data()];
    bk.resize(G.n);
    d2h(e, bk.data(), pkeep.p, G.n);
    const size_t c2 = std::max(nb2, 1);
    e->tmp_ij.reserve(c2); e->tmp_a.reserve(c2); e->tmp_b.reserve(c2); e->tmp_c.reserve(c2);
    e->tmp_rank.reserve(c2);
    launch_compact_bonds(e->stream, G.n, pkeep.p, poff.p, rij.p, G.k4.p, G.mx4.p, G.D4.p, G.rank.p,
                         e->tmp_ij.p, e->tmp_a.p, e->tmp_b.p, e->tmp_c.p, e->tmp_rank.p);
    HIPCHK(hipGetLastError());
    sync(e);
    G.ij.swap(e->tmp_ij); G.k4.swap(e->tmp_a); G.mx4.swap(e->tmp_b); G.D4.swap(e->tmp_c);
    G.rank.swap(e->tmp_rank);
    int w = 0;
    for (int b = 0; b < G.n; b++)
      if (bk[b]) { G.hid1[w] = G.hid1[b]; G.hid2[w] = G.hid2[b]; w++; }
    G.hid1.resize(w);
    G.hid2.resize(w);
    G.n = nb2;
  }
  {
    int off = 0;
    for (int gi : e->bord) {
      e->bg[gi].bond_off = off;
      off += e->bg[gi].n;
    }
  }
  if (!pend.empty()) {
    // new index of a kept bond = its group's new offset + kept bonds before it
    std::vector<std::vector<int>> rank(e->bg.size());
    for (size_t g = 0; g < e->bg.size(); g++) {
      rank[g].resize(bkeep[g].size());
      int r = 0;
      for (size_t b = 0; b < bkeep[g].size(); b++) {
        rank[g][b] = bkeep[g][b] ? r : -1;
        r += bkeep[g][b];
      }
    }
    std::vector<int> kept;
    for (int x : pend)
      for (size_t g = 0; g < e->bg.size(); g++) {
        if (x < old_off[g] || x >= old_off[g] + old_n[g]) continue;
        const int r = rank[g][x - old_off[g]];
        if (r >= 0) kept.push_back(e->bg[g].bond_off + r);
        break;
      }
    if (!kept.empty()) h2d(e, broken_list_of(e, e->t), kept.data(), kept.size());
    const int nk2 = (int)kept.size();
    h2d(e, &e->ring.p[ring_i].nbroken, &nk2, 1);
    sync(e);
  }
  for (auto& G : e->dg) {
    if (G.n <= 0) continue;
    const bool hist = G.kind == 3;
    pkeep.reserve((size_t)G.n + 1);
    poff.reserve((size_t)G.n + 1);
    rij.reserve(G.n);
    launch_remap_pairs(e->stream, G.n, G.ij.p, nslot.p, rij.p, pkeep.p);
    const int nd = scan_total(e, pkeep.p, poff.p, G.n);
    e->tmp_ij.reserve((size_t)nd + ESP_APP_CAP);
    e->tmp_key.reserve((size_t)nd + ESP_APP_CAP);
    if (hist) e->tmp_ff.reserve((size_t)nd + ESP_APP_CAP);
    launch_pack_dyn(e->stream, G.n, pkeep.p, poff.p, rij.p, hist ? G.ff.p : nullptr, e->tmp_ij.p,
                    e->tmp_key.p, hist ? e->tmp_ff.p : nullptr);
    HIPCHK(hipGetLastError());
    sync(e);
    G.ij.swap(e->tmp_ij);
    G.key.swap(e->tmp_key);
    if (hist) G.ff.swap(e->tmp_ff);
    G.n = nd;
  }
  // the flagged-particle list of the device bookkeeping
  UpdState st;
  d2h(e, &st, e->upd.p, 1);
  sync(e);
  if (st.fcount > 0) {
    std::vector<int> fl(st.fcount);
    d2h(e, fl.data(), e->flist.p, st.fcount);
    sync(e);
    int w = 0;
    for (int v : fl)
      if (v >= 0 && v < n && hslot[v] >= 0) fl[w++] = hslot[v];
    h2d(e, e->flist.p, fl.data(), w);
    h2d(e, &e->upd.p->fcount, &w, 1);
    sync(e);
  }
  // per-particle arrays
  compact_buf(e, e->pos, n, nk, 1);
  compact_buf(e, e->vel, n, nk, 1);
  compact_buf(e, e->ang, n, nk, 1);
  compact_buf(e, e->quat, n, nk, 1);
  compact_buf(e, e->old, n, nk, 1);
  compact_buf(e, e->frc, n, nk, 1);
  compact_buf(e, e->mom, n, nk, 1);
  compact_buf(e, e->flags, n, nk, 1);
  compact_buf(e, e->id, n, nk, 1);
  compact_buf(e, e->inert, n, nk, 1);
  compact_buf(e, e->init, n, nk, 1);
  compact_buf(e, e->tag, n, nk, 1);
  // host mirrors (per slot)
  compact_host(e->h_id, hkeep, 1);
  compact_host(e->h_tag, hkeep, 1);
  compact_host(e->h_pos, hkeep, 3);
  compact_host(e->h_vel, hkeep, 3);
  compact_host(e->h_rad, hkeep, 1);
  compact_host(e->h_mass, hkeep, 1);
  compact_host(e->h_flags, hkeep, 1);
  compact_host(e->h_init, hkeep, 3);
  compact_host(e->h_quat, hkeep, 4);
  compact_host(e->h_ang, hkeep, 3);
  compact_host(e->h_inert, hkeep, 1);
  compact_host(e->h_order, hkeep, 1);
  e->ord_valid = false;
  e->id2slot.clear();
  for (int i = 0; i < nk; i++) e->id2slot[e->h_id[i]] = i;
  e->n = nk;
  for (auto& G : e->bg) rebuild_bond_hash(e, G);
  for (auto& G : e->dg) rebuild_presence(e, G);
  sync(e);
}

// First part of the table build: positions wrapped into the circular x range,
// each particle's cell (-1 outside the padded grid), the displacements reset;
// returns the number of particles outside.  Reads and writes no interaction
// storage, so the host step runs it before the bond-break update and folds.
int rebuild_count(esp_engine* e)
{
  PView p = e->view();
  // circular x: particles that crossed an x end take their image position
  // (CParticle::setCircular when the ghost copy takes over at the table rebuild)
  if (e->grid.circ_x) launch_wrap_x(e->stream, p, e->grid.x0, e->grid.period, e->init.p);
  zero(e, e->cell_count.p, e->ncell + 1);
  zero(e, e->dev_int.p, 1);
  launch_cells_count(e->stream, p, e->grid, e->cell.p, e->cell_count.p, e->dev_int.p);
  int nout = 0;
  d2h(e, &nout, e->dev_int.p, 1);
  sync(e);
  return nout;
}

// The rest of the table build after rebuild_count (nout particles outside),
// with nothing moved in between: erasure, cell lists, the full candidate list.
void rebuild_rest(esp_engine* e, long long step, int nout)
{
  PView p = e->view();
  if (nout > 0) {
    erase_outside(e);
    p = e->view();
    zero(e, e->cell_count.p, e->ncell + 1);
    zero(e, e->dev_int.p, 1);
    launch_cells_count(e->stream, p, e->grid, e->cell.p, e->cell_count.p, e->dev_int.p);
    HIPCHK(hipGetLastError());
  }
  scan_total(e, e->cell_count.p, e->cell_start.p, e->ncell);
  zero(e, e->cell_fill.p, e->ncell);
  launch_cells_fill(e->stream, e->n, e->cell.p, e->cell_start.p, e->cell_fill.p, e->cell_items.p);
  launch_cells_sort(e->stream, e->ncell, e->cell_start.p, e->cell_items.p);
  // full candidate list (getFullList), sorted by (id1, id2)
  launch_pairs(e->stream, 0, 0, p, e->grid, e->alpha, e->cell.p, e->cell_start.p,
               e->cell_items.p, e->cand_cnt.p, nullptr, nullptr, nullptr, nullptr);
  const int total = scan_total(e, e->cand_cnt.p, e->cand_off.p, e->n);
  e->cand_p.reserve(std::max(total, 1));
  e->cand_q.reserve(std::max(total, 1));
  e->cand_key.reserve(std::max(total, 1));
  launch_pairs(e->stream, 1, 0, p, e->grid, e->alpha, e->cell.p, e->cell_start.p,
               e->cell_items.p, nullptr, e->cand_off.p, e->cand_p.p, e->cand_q.p, e->cand_key.p);
  HIPCHK(hipGetLastError());
  e->ncand = total;
  e->timestamp++;
  e->rebuild_steps.push_back((int)step);
}

void rebuild(esp_engine* e, long long step) { rebuild_rest(e, step, rebuild_count(e)); }

// flagged pair list (getNewList) and clearing of all flags
void new_pair_list(esp_engine* e)
{
  PView p = e->view();
  launch_pairs(e->stream, 0, 1, p, e->grid, e->alpha, e->cell.p, e->cell_start.p,
               e->cell_items.p, e->cand_cnt.p, nullptr, nullptr, nullptr, nullptr);
  const int total = scan_total(e, e->cand_cnt.p, e->cand_off.p, e->n);
  e->np_p.reserve(std::max(total, 1));
  e->np_q.reserve(std::max(total, 1));
  e->np_key.reserve(std::max(total, 1));
  launch_pairs(e->stream, 1, 1, p, e->grid, e->alpha, e->cell.p, e->cell_start.p,
               e->cell_items.p, nullptr, e->cand_off.p, e->np_p.p, e->np_q.p, e->np_key.p);
  launch_clear_flag_bit(e->stream, p, e->upd.p);
  HIPCHK(hipGetLastError());
  e->nnew = total;
}

int dyn_appended(esp_engine* e, DynGroup& G);
long long id_key(esp_engine* e, long long skey);
void dyn_view(esp_engine* e, DynGroup& G, std::vector<long long>& key, std::vector<double4>* ff);

// Pair events of a tracked group, part 1: appendix pairs that queued steps
// created on the device since the last harvest, logged with the step of their
// creation (the device writes it beside each appendix entry).  Reads only, so
// queries may call it at any time; the slots must be the ones the appendix
// was written with (no erasure in between).
void harvest_appendix(esp_engine* e, DynGroup& G)
{
  if (!G.track || !e->started) return;
  const int na = dyn_appended(e, G);
  const int m = na - G.app_logged;
  if (m <= 0) return;
  std::vector<long long> k(m);
  std::vector<int> st(m);
  d2h(e, k.data(), G.key.p + G.n + G.app_logged, m);
  d2h(e, st.data(), G.app_step.p + G.app_logged, m);
  sync(e);
  for (int i = 0; i < m; i++) {
    const long long ik = id_key(e, k[i]);
    G.ev.push_back(make_int4(st[i], (int)(ik >> 32), (int)(ik & 0xffffffffLL), 1));
    G.ev_pend.push_back(ik);
  }
  G.app_logged = na;
}

// Pair events of a tracked group, part 2: after a host-side interaction update
// (table rebuild with erasure, flagged-list update) at `step`, the stored pairs
// as sorted id keys against those of the last logged update plus the appendix
// pairs logged since: a key whose count differs gives one event per unit of
// difference, kind +1 (created) or -1 (removed).
void log_pair_diff(esp_engine* e, DynGroup& G, long long step)
{
  std::vector<long long> cur;
  dyn_view(e, G, cur, nullptr);
  std::vector<long long>& prev = G.ev_prev;
  if (!G.ev_pend.empty()) {
    prev.insert(prev.end(), G.ev_pend.begin(), G.ev_pend.end());
    std::sort(prev.begin(), prev.end());
    G.ev_pend.clear();
  }
  const size_t np = prev.size(), nc = cur.size();
  size_t a = 0, b = 0;
  while (a < np || b < nc) {
    long long k;
    int kind;
    if (b == nc || (a < np && prev[a] < cur[b])) {
      k = prev[a++];
      kind = -1;
    } else if (a == np || cur[b] < prev[a]) {
      k = cur[b++];
      kind = 1;
    } else {
      a++;
      b++;
      continue;
    }
    G.ev.push_back(make_int4((int)step, (int)(k >> 32), (int)(k & 0xffffffffLL), kind));
  }
  prev.swap(cur);
}

// Break events logged on the device move to the host log.  Touches nothing the
// steps read, so queries may call it at any time.
void fold_events(esp_engine* e)
{
  UpdState st;
  d2h(e, &st, e->upd.p, 1);
  sync(e);
  if (st.evt_count <= 0) return;
  std::vector<int4> ev(st.evt_count);
  d2h(e, ev.data(), e->evt.p, st.evt_count);
  const int zero_int = 0;
  h2d(e, &e->upd.p->evt_count, &zero_int, 1);
  sync(e);
  for (const int4& v : ev) {
    e->breaks.push_back(v.x);
    e->breaks.push_back(v.y);
    e->breaks.push_back(v.z);
  }
  e->dead_pending += st.evt_count;
}

// Device bookkeeping folded back into plain lists: dead bonds compacted out,
// each dynamic group's appendix merged into its key-sorted main list.  Record
// indices change, so this runs only in host bookkeeping, right before the
// force-record layout and the CSR are rebuilt (rebuild_csr).
void fold_appendices(esp_engine* e);

void fold(esp_engine* e)
{
  fold_events(e);
  if (e->dead_pending > 0) {
    e->dead_pending = 0;
    // compact dead bonds
    int off = 0;
    for (int gi : e->bord) {
      BondGroup& G = e->bg[gi];
      if (G.n > 0) {
        e->keep.reserve(G.n + 1);
        e->keep_off.reserve(G.n + 1);
        launch_alive_flags(e->stream, G.n, G.ij.p, e->keep.p);
        const int nb2 = scan_total(e, e->keep.p, e->keep_off.p, G.n);
        if (nb2 != G.n) {
          std::vector<int> keep(G.n);
          d2h(e, keep.data(), e->keep.p, G.n);
          const size_t c2 = std::max(nb2, 1);
          e->tmp_ij.reserve(c2); e->tmp_a.reserve(c2); e->tmp_b.reserve(c2); e->tmp_c.reserve(c2);
          e->tmp_rank.reserve(c2);
          launch_compact_bonds(e->stream, G.n, e->keep.p, e->keep_off.p, G.ij.p, G.k4.p, G.mx4.p,
                               G.D4.p, G.rank.p, e->tmp_ij.p, e->tmp_a.p, e->tmp_b.p, e->tmp_c.p,
                               e->tmp_rank.p);
          HIPCHK(hipGetLastError());
          sync(e);
          G.ij.swap(e->tmp_ij); G.k4.swap(e->tmp_a); G.mx4.swap(e->tmp_b); G.D4.swap(e->tmp_c);
          G.rank.swap(e->tmp_rank);
          int w = 0;
          for (int b = 0; b < G.n; b++)
            if (keep[b]) { G.hid1[w] = G.hid1[b]; G.hid2[w] = G.hid2[b]; w++; }
          G.hid1.resize(w);
          G.hid2.resize(w);
          G.n = nb2;
          rebuild_bond_hash(e, G);  // drops the erased slots
        }
      }
      G.bond_off = off;
      off += G.n;
    }
  }
  fold_appendices(e);
}

// Each dynamic group's appendix merged into its key-sorted main list (record
// indices of the group change; bond indices do not).
void fold_appendices(esp_engine* e)
{
  for (auto& G : e->dg) harvest_appendix(e, G);
  std::vector<int> napp(e->dg.size(), 0);
  for (size_t g = 0; g < e->dg.size(); g++) d2h(e, &napp[g], e->dg[g].napp.p, 1);
  sync(e);
  for (size_t g = 0; g < e->dg.size(); g++) {
    DynGroup& G = e->dg[g];
    const int na = napp[g];
    if (na <= 0) continue;
    const bool hist = G.kind == 3;
    // appendix: batches in key order, not sorted overall -> sort it on the host
    std::vector<int2> aij(na);
    std::vector<long long> akey(na);
    std::vector<double4> aff(hist ? na : 0);
    d2h(e, aij.data(), G.ij.p + G.n, na);
    d2h(e, akey.data(), G.key.p + G.n, na);
    if (hist) d2h(e, aff.data(), G.ff.p + G.n, na);
    sync(e);
    std::vector<int> perm(na);
    for (int i = 0; i < na; i++) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return akey[a] < akey[b]; });
    std::vector<int2> sij(na);
    std::vector<long long> skey(na);
    std::vector<double4> sff(hist ? na : 0);
    for (int i = 0; i < na; i++) {
      sij[i] = aij[perm[i]];
      skey[i] = akey[perm[i]];
      if (hist) sff[i] = aff[perm[i]];
    }
    e->tmp_ij2.reserve(na);
    e->tmp_key2.reserve(na);
    e->tmp_ff.reserve(na);
    h2d(e, e->tmp_ij2.p, sij.data(), na);
    h2d(e, e->tmp_key2.p, skey.data(), na);
    if (hist) h2d(e, e->tmp_ff.p, sff.data(), na);
    const int tot = G.n + na;
    e->tmp_ij.reserve((size_t)tot + ESP_APP_CAP);
    e->tmp_key.reserve((size_t)tot + ESP_APP_CAP);
    e->tmp_a.reserve((size_t)tot + ESP_APP_CAP);
    launch_merge(e->stream, G.n, G.ij.p, G.key.p, hist ? G.ff.p : nullptr, na, e->tmp_ij2.p,
                 e->tmp_key2.p, hist ? e->tmp_ff.p : nullptr, e->tmp_ij.p, e->tmp_key.p,
                 hist ? e->tmp_a.p : nullptr);
    HIPCHK(hipGetLastError());
    sync(e);
    G.ij.swap(e->tmp_ij);
    G.key.swap(e->tmp_key);
    if (hist) G.ff.swap(e->tmp_a);
    G.n = tot;
    zero(e, G.napp.p, 2);
    G.app_logged = 0;
    reserve_group(G, G.n, hist);  // keeps content when the capacity suffices
    rebuild_presence(e, G);       // room for the next appendix
  }
  sync(e);
}

// PIS_NE::update / PIS_ED::update (pi_storage_ne.hpp:49-102, pi_storage_ed.hpp:84-161)
// on the host bookkeeping path; the group's appendix is empty here (folded).
void dyn_update(esp_engine* e, DynGroup& G, bool& new_list_taken)
{
  const bool hist = G.kind == 3;
  PView p = e->view();
  ExclSet ex = excl_set(e, G);
  if (G.update_ts != e->timestamp) {
    const int nc = e->ncand;
    e->keep.reserve(std::max(nc, 1) + 1);
    e->keep_off.reserve(std::max(nc, 1) + 1);
    if (hist) {
      e->src.reserve(std::max(nc, 1));
      launch_ed_mark(e->stream, nc, e->cand_p.p, e->cand_q.p, e->cand_key.p, ex, G.key.p, G.n,
                     G.ij.p, p, e->grid.period, e->keep.p, e->src.p);
    } else {
      launch_ne_mark(e->stream, nc, e->cand_p.p, e->cand_q.p, e->id.p, ex, e->keep.p);
    }
    const int total = scan_total(e, e->keep.p, e->keep_off.p, nc);
    e->tmp_ij.reserve((size_t)total + ESP_APP_CAP);
    e->tmp_key.reserve((size_t)total + ESP_APP_CAP);
    if (hist) e->tmp_ff.reserve((size_t)total + ESP_APP_CAP);
    launch_pair_emit(e->stream, nc, e->keep.p, e->keep_off.p, e->cand_p.p, e->cand_q.p,
                     e->cand_key.p, hist ? e->src.p : nullptr, hist ? G.ff.p : nullptr,
                     e->tmp_ij.p, e->tmp_key.p, hist ? e->tmp_ff.p : nullptr);
    HIPCHK(hipGetLastError());
    G.ij.swap(e->tmp_ij);
    G.key.swap(e->tmp_key);
    if (hist) G.ff.swap(e->tmp_ff);
    G.n = total;
    rebuild_presence(e, G);
  } else {
    int nc = 0;
    if (!new_list_taken) {
      new_pair_list(e);
      new_list_taken = true;
      nc = e->nnew;
    }
    if (nc > 0) {
      e->keep.reserve(nc + 1);
      e->keep_off.reserve(nc + 1);
      launch_filter_new(e->stream, nc, e->np_p.p, e->np_q.p, e->id.p, e->np_key.p, ex,
                        HashRef{G.hs.p, G.hs_mask}, hist ? 1 : 0, e->keep.p);
      const int add = scan_total(e, e->keep.p, e->keep_off.p, nc);
      if (add > 0) {
        e->tmp_ij2.reserve(add);
        e->tmp_key2.reserve(add);
        launch_pack_pairs(e->stream, nc, e->keep.p, e->keep_off.p, e->np_p.p, e->np_q.p,
                          e->np_key.p, e->tmp_ij2.p, e->tmp_key2.p);
        const int tot = G.n + add;
        e->tmp_ij.reserve((size_t)tot + ESP_APP_CAP);
        e->tmp_key.reserve((size_t)tot + ESP_APP_CAP);
        if (hist) e->tmp_ff.reserve((size_t)tot + ESP_APP_CAP);
        launch_merge(e->stream, G.n, G.ij.p, G.key.p, hist ? G.ff.p : nullptr, add, e->tmp_ij2.p,
                     e->tmp_key2.p, nullptr, e->tmp_ij.p, e->tmp_key.p,
                     hist ? e->tmp_ff.p : nullptr);
        HIPCHK(hipGetLastError());
        sync(e);
        G.ij.swap(e->tmp_ij);
        G.key.swap(e->tmp_key);
        if (hist) G.ff.swap(e->tmp_ff);
        G.n = tot;
        rebuild_presence(e, G);
      }
    }
  }
  G.update_ts = e->timestamp;
}

void wall_params(esp_engine* e, WallParam* out)
{
  for (int k = 0; k < ESP_MAX_WALL_GROUPS; k++) memset(&out[k], 0, sizeof(WallParam));
  for (size_t k = 0; k < e->word.size(); k++) {
    const WallGroup& G = e->wg[e->word[k]];
    const Wall& W = e->walls[G.wall];
    for (int c = 0; c < 3; c++) {
      out[k].o[c] = W.o[c];
      out[k].nrm[c] = W.n[c];
      out[k].disp[c] = W.o[c] - W.old[c];
    }
    out[k].k = G.k;
    out[k].bonded = G.bonded;
    out[k].tagged = G.tagged;
  }
}

// force records: main lists of the dynamic groups (name order), bonded groups
// (name order), then one appendix region per dynamic group; the CSR covers the
// first two, the per-particle overflow slots the appendices
void rebuild_csr(esp_engine* e)
{
  long long rec = 0;
  for (int gi : e->dord) { e->dg[gi].rec_off = (int)rec; rec += e->dg[gi].n; }
  for (int gi : e->bord) { e->bg[gi].rec_off = (int)rec; rec += e->bg[gi].n; }
  for (int gi : e->dord) { e->dg[gi].app_rec_off = (int)rec; rec += ESP_APP_CAP; }
  if (rec >= (1ll << 29)) throw EspError(ESP_ERR_CAPACITY, "more than 2^29 interactions");
  for (int gi : e->dord)
    if (e->dg[gi].n >= (1 << 27)) throw EspError(ESP_ERR_CAPACITY, "more than 2^27 pairs in one group");
  e->nrec = (int)rec;
  e->slot_of.reserve((size_t)std::max<long long>(rec, 1));
  zero(e, e->inc_cnt.p, e->n + 1);
  for (int gi : e->dord) launch_inc_count(e->stream, e->dg[gi].n, e->dg[gi].ij.p, e->inc_cnt.p);
  for (int gi : e->bord) launch_inc_count(e->stream, e->bg[gi].n, e->bg[gi].ij.p, e->inc_cnt.p);
  const int total = scan_total(e, e->inc_cnt.p, e->inc_off.p, e->n);
  e->inc_ent.reserve(std::max(total, 1));
  e->inc_ent64.reserve(std::max(total, 1));
  zero(e, e->inc_fill.p, e->n + 1);
  // summation order: dynamic groups by name (storage order), then bonded groups
  // by name (connection order)
  int grp = 0;
  for (int gi : e->dord) {
    DynGroup& G = e->dg[gi];
    launch_inc_fill(e->stream, G.n, G.ij.p, G.rec_off, G.kind == 3 ? 1 : 0, nullptr, grp++,
                    e->inc_off.p, e->inc_fill.p, e->inc_ent64.p);
  }
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    launch_inc_fill(e->stream, G.n, G.ij.p, G.rec_off, 1, G.rank.p, grp++, e->inc_off.p,
                    e->inc_fill.p, e->inc_ent64.p);
  }
  launch_inc_sort(e->stream, e->n, e->inc_off.p, e->inc_ent64.p, e->inc_ent.p);
  launch_slot_of(e->stream, total, e->inc_ent.p, e->slot_of.p);
  // slots start at zero: a contact that is not touching keeps zero slots
  // without rewriting them every step
  e->xbase = total;
  const size_t nslot = (size_t)total + (size_t)std::max(e->n, 1) * ESP_XSLOTS;
  e->slots.reserve(nslot * e->ss);
  zero(e, e->slots.p, nslot * e->ss);
  e->sact.reserve(nslot);
  zero(e, e->sact.p, nslot);
  zero(e, e->xcnt.p, std::max(e->n, 1));
  HIPCHK(hipGetLastError());
}

BondSet bond_set(esp_engine* e)
{
  BondSet bs;
  memset(&bs, 0, sizeof bs);
  bs.ss = e->ss;
  bs.slots = e->slots.p;
  bs.sact = e->sact.p;
  bs.slot_of = e->slot_of.p;
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    const int k = bs.count++;
    bs.off[k] = G.bond_off;
    bs.n[k] = G.n;
    bs.rec_off[k] = G.rec_off;
    bs.ij[k] = G.ij.p;
    bs.hs[k] = HashRef{G.hs.p, G.hs_mask};
  }
  return bs;
}

// ---------------------------------------------------------------------------
// one step's kernels
// ---------------------------------------------------------------------------
// segments of the fused force-law launch: bonded groups, then dynamic groups
int* broken_list_of(esp_engine* e, long long step)
{
  return e->broken_list.p + (size_t)(step & 1) * e->broken_cap;
}

// flip: this step reads the second history buffer (odd steps of a batch)
ForceLaunch force_launch(esp_engine* e, long long step, bool flip)
{
  ForceLaunch fl;
  memset(&fl, 0, sizeof fl);
  fl.do2d = e->do2d;
  fl.period = e->grid.period;
  fl.ss = e->ss;
  fl.dt = e->dt;
  fl.step = (double)step;
  fl.slots = e->slots.p;
  fl.sact = e->sact.p;
  fl.slot_of = e->slot_of.p;
  fl.broken_list = broken_list_of(e, step);
  fl.crit = e->track_crit ? e->crit.p : nullptr;
  int blk = 0;
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    if (G.n <= 0) continue;
    ForceSeg& s = fl.seg[fl.count++];
    s.kind = ESP_SEG_ROT_BOND;
    s.blk0 = blk;
    s.n = G.n;
    s.rec_off = G.rec_off;
    s.bond_off = G.bond_off;
    s.ij = G.ij.p;
    s.k4 = G.k4.p;
    s.mx4 = G.mx4.p;
    s.D4 = G.D4.p;
    s.truncated = G.prm.truncated;
    s.beta1 = G.prm.beta1;
    s.beta2 = G.prm.beta2;
    s.sc = G.sc;
    blk += force_blocks(G.n);
  }
  for (int gi : e->dord) {
    DynGroup& G = e->dg[gi];
    ForceSeg& s = fl.seg[fl.count++];
    s.kind = G.kind == 1 ? ESP_SEG_NROT_ELASTIC : G.kind == 2 ? ESP_SEG_ROT_ELASTIC
                                                              : ESP_SEG_ROT_FRICTION;
    s.blk0 = blk;
    s.n = G.n;
    s.rec_off = G.rec_off;
    s.app_rec_off = G.app_rec_off;
    s.n_app = G.napp.p + 1;
    s.ij = G.ij.p;
    if (G.kind == 3) {
      s.ff = flip ? G.ff2.p : G.ff.p;
      s.ff_out = flip ? G.ff.p : G.ff2.p;
    }
    s.k = G.k;
    s.ks = G.ks;
    s.mu_d = G.mu_d;
    s.mu_s = G.mu_s;
    s.rigid = G.rigid;
    s.scaling = G.scaling;
    s.meanR = G.meanR;
    blk += force_blocks(G.n + ESP_APP_CAP);
  }
  fl.nblocks = blk;
  return fl;
}

ScLaunch sc_launch(esp_engine* e, long long step)
{
  ScLaunch sl;
  memset(&sl, 0, sizeof sl);
  sl.sig = e->sig.p;
  sl.period = e->grid.period;
  sl.broken_list = broken_list_of(e, step);
  sl.step = (double)step;
  int blk = 0;
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    if (!G.sc || G.n <= 0) continue;
    const int k = sl.count++;
    sl.blk0[k] = blk;
    sl.n[k] = G.n;
    sl.bond_off[k] = G.bond_off;
    sl.ij[k] = G.ij.p;
    sl.D4[k] = G.D4.p;
    sl.sigma3[k] = G.sc_sigma3;
    sl.cohesion[k] = G.sc_cohesion;
    sl.tan_angle[k] = G.sc_tan;
    blk += force_blocks(G.n);
  }
  sl.nblocks = blk;
  return sl;
}
// last: the host may read the particles' total forces and moments after this
// step (the last step of a queued batch, every host step); other steps keep
// them in registers only
void launch_step(esp_engine* e, const StepCtl& ctl, long long step, bool flip, bool last)
{
  PView p = e->view();
  const ForceLaunch fl = force_launch(e, step, flip);
  UpdParams up;
  const bool dev_update = ctl.guarded && !e->bg.empty();
  if (dev_update) {
    // bookkeeping of the previous step's bond breaks, on the device
    memset(&up, 0, sizeof up);
    up.bs = bond_set(e);
    up.broken_prev = broken_list_of(e, step - 1);
    up.evt = e->evt.p;
    up.step = step;
    up.g = e->grid;
    up.cell = e->cell.p;
    up.cell_start = e->cell_start.p;
    up.cell_items = e->cell_items.p;
    up.st = e->upd.p;
    up.flist = e->flist.p;
    up.xcnt = e->xcnt.p;
    if (!e->dg.empty()) {
      DynGroup& G = e->dg[e->dord[0]];
      DynAppend& da = up.da;
      da.ij = G.ij.p;
      da.key = G.key.p;
      da.ff = G.kind == 3 ? (flip ? G.ff2.p : G.ff.p) : nullptr;
      da.n_app = G.napp.p;
      da.n_main = G.n;
      da.rec_off = G.rec_off;
      da.app_rec_off = G.app_rec_off;
      da.is_hist = G.kind == 3;
      da.present = HashRef{G.hs.p, G.hs_mask};
      da.ex = excl_set(e, G);
      da.slot_of = e->slot_of.p;
      da.xbase = e->xbase;
      up.has_group = 1;
      up.useg = 0;
      while (up.useg < fl.count && fl.seg[up.useg].kind == ESP_SEG_ROT_BOND) up.useg++;
    }
  }
  {
    // the pair-event log stamps the step of each pair the device appends
    ForceExt ext;
    ext.app_step = nullptr;
    if (dev_update && !e->dg.empty() && e->dg[e->dord[0]].track)
      ext.app_step = e->dg[e->dord[0]].app_step.p;
    const LaunchEvents t = launch_events(e);
    note_launch(e, t,
                launch_forces(e->stream, ctl, p, fl, dev_update ? &up : nullptr, t.a, t.b, &ext),
                ESP_KCLASS_FORCES, ctl.step_id);
  }
  GatherParams gp;
  memset(&gp, 0, sizeof gp);
  gp.inc_off = e->inc_off.p;
  gp.xcnt = e->xcnt.p;
  gp.xbase = e->xbase;
  gp.grid = e->grid;
  gp.slots = e->slots.p;
  gp.sact = e->sact.p;
  gp.dt = e->dt;
  gp.rot = e->rot;
  gp.do2d = e->do2d;
  gp.nwg = (int)e->wg.size();
  wall_params(e, gp.wall);
  gp.ngrav = (int)e->gord.size();
  for (size_t k = 0; k < e->gord.size(); k++)
    for (int c = 0; c < 3; c++) gp.g[k][c] = e->grav[e->gord[k]].g[c];
  gp.wall_part = e->wall_part.p + (size_t)(ctl.step_id - 1) * gather_blocks(e->n) * gp.nwg * 3;
  gp.sig = e->ss == 9 ? e->sig.p : nullptr;
  gp.n_app = e->dg.empty() ? nullptr : e->dg[e->dord[0]].napp.p;
  gp.store_fm = last ? 1 : 0;
  gp.ndamp = (int)e->damp.size();
  for (int d = 0; d < gp.ndamp; d++) {
    gp.damp_type[d] = e->damp[d].type;
    gp.damp_visc[d] = e->damp[d].visc;
    gp.damp_maxit[d] = e->damp[d].maxit;
  }
  gp.inert = e->inert.p;
  gp.init = e->init.p;
  {
    const LaunchEvents t = launch_events(e);
    note_launch(e, t, launch_gather_integrate(e->stream, ctl, p, gp, t.a, t.b), ESP_KCLASS_GATHER,
                ctl.step_id);
  }
  if (e->ss == 9) launch_sc_break(e->stream, ctl, p, sc_launch(e, step));
  HIPCHK(hipGetLastError());
}

unsigned long long thresh_bits(esp_engine* e)
{
  const double a = 0.5 * e->alpha;
  const double sr = a * a;
  unsigned long long b;
  memcpy(&b, &sr, 8);
  return b;
}

void record_walls(esp_engine* e, const std::vector<double>& h, int steps)
{
  const int nwg = (int)e->wg.size();
  for (int s = 0; s < steps; s++)
    for (int w = 0; w < (int)e->walls.size(); w++) {
      double f[3] = {0, 0, 0};
      // a wall's force is the sum over its wall groups (each adds to the same CWall)
      for (int k = 0; k < nwg; k++)
        if (e->wg[e->word[k]].wall == w)
          for (int c = 0; c < 3; c++) f[c] = f[c] + h[((size_t)s * nwg + k) * 3 + c];
      for (int c = 0; c < 3; c++) {
        e->walls[w].hist.push_back(f[c]);
        e->walls[w].force[c] = f[c];
      }
    }
}

void apply_wall_motion(esp_engine* e)
{
  for (auto& W : e->walls)
    if (W.has_motion)
      for (int c = 0; c < 3; c++) W.o[c] = W.o[c] + W.step[c];
}

StepCtl make_ctl(esp_engine* e, long long step, int step_id, int guarded)
{
  StepCtl ctl;
  ctl.prev = e->ring.p + ((step - 1) % e->ring_size);
  ctl.cur = e->ring.p + (step % e->ring_size);
  ctl.halted = e->halted.p;
  ctl.thresh = thresh_bits(e);
  ctl.step_id = step_id;
  ctl.guarded = guarded;
  ctl.reason = &e->upd.p->halt_reason;
  return ctl;
}

// Shear histories: a step reads G.ff and writes G.ff2, then the two swap.
// After host bookkeeping rewrote G.ff both buffers are made equal, so flag bit
// 2 ("zero in both buffers") holds for the next step.
void sync_history_buffers(esp_engine* e)
{
  for (auto& G : e->dg) {
    if (G.kind != 3) continue;
    G.ff2.reserve(G.ff.cap);
    if (G.n > 0)
      HIPCHK(hipMemcpyAsync(G.ff2.p, G.ff.p, (size_t)G.n * sizeof(double4),
                            hipMemcpyDeviceToDevice, e->stream));
  }
}

void swap_history_buffers(esp_engine* e)
{
  for (auto& G : e->dg)
    if (G.kind == 3) G.ff.swap(G.ff2);
}

// one step with host bookkeeping: the first step, a rebuild, or an
// incremental update that did not fit the device appendix
void host_step(esp_engine* e, int reason)
{
  e->num_host_phases++;
  const long long step = e->t + 1;
  StepStatus prev;
  memset(&prev, 0, sizeof prev);
  if (!e->first_time) {
    d2h(e, &prev, e->ring.p + ((step - 1) % e->ring_size), 1);
    sync(e);
  }
  // forced search of the first runOneStep, or displacement past the threshold;
  // after a checkpoint's search displacements count from there
  const bool search = e->first_time || (!e->forced_search && prev.maxdisp2 > thresh_bits(e));
  // The reference searches first: particles that left the grid are erased with
  // their bonds before the bonds are tested.  The search's cell pass runs
  // first, so the break update knows whether any particle is outside.
  const int nout = search ? rebuild_count(e) : 0;
  // PIS_EB::update for the previous step's breaks (already applied on the
  // device when the queue stopped in the incremental update); bonds of
  // particles the search erases are skipped
  if (prev.nbroken > 0 && reason != 2 && !e->bg.empty()) {
    StepCtl c0 = make_ctl(e, step, 1, 0);
    launch_break_apply(e->stream, c0, e->view(), bond_set(e), broken_list_of(e, step - 1),
                       e->upd.p, e->flist.p, e->evt.p, step, nout > 0 ? &e->grid : nullptr);
    HIPCHK(hipGetLastError());
  }
  fold(e);
  if (search) rebuild_rest(e, step, nout);
  e->first_time = false;
  e->forced_search = false;
  e->t = step;
  apply_wall_motion(e);
  // TSubLattice::updateInteractions (SubLattice.hpp:2416-2466)
  const bool need_update = prev.nbroken > 0;
  if (e->timestamp > e->last_ns || need_update) {
    bool new_list_taken = false;
    for (int gi : e->dord) dyn_update(e, e->dg[gi], new_list_taken);
    for (auto& G : e->dg)
      if (G.track) log_pair_diff(e, G, step);
    if (!e->wg.empty()) {
      WallParam w[ESP_MAX_WALL_GROUPS];
      wall_params(e, w);
      launch_wall_update(e->stream, e->view(), e->grid, (int)e->wg.size(), w);
      HIPCHK(hipGetLastError());
    }
    e->last_ns = e->timestamp;
  }
  rebuild_csr(e);
  sync_history_buffers(e);
  launch_clear_status(e->stream, e->ring.p, e->ring_size, step, 1, e->halted.p);
  launch_step(e, make_ctl(e, step, 1, 0), step, false, true);
  swap_history_buffers(e);
  const int nwg = (int)e->wg.size();
  launch_wall_finalize(e->stream, e->wall_part.p, gather_blocks(e->n), nwg, 1, e->wall_hist.p);
  std::vector<double> h((size_t)nwg * 3);
  if (nwg) d2h(e, h.data(), e->wall_hist.p, h.size());
  sync(e);
  collect_timing(e, 1);
  if (nwg) record_walls(e, h, 1);
}

// up to K steps queued without host bookkeeping; returns steps completed
int fast_batch(esp_engine* e, int k, int* reason)
{
  const long long step0 = e->t + 1;
  const int nwg = (int)e->wg.size();
  std::vector<double> before;
  for (auto& W : e->walls) before.insert(before.end(), W.o, W.o + 3);
  std::vector<std::vector<double>> wall_o;
  launch_clear_status(e->stream, e->ring.p, e->ring_size, step0, k, e->halted.p);
  for (int s = 0; s < k; s++) {
    apply_wall_motion(e);  // pre-time-step runnables of this step
    std::vector<double> snap;
    for (auto& W : e->walls) snap.insert(snap.end(), W.o, W.o + 3);
    wall_o.push_back(snap);
    const long long step = step0 + s;
    launch_step(e, make_ctl(e, step, s + 1, 1), step, (s & 1) != 0, s == k - 1);
  }
  launch_wall_finalize(e->stream, e->wall_part.p, gather_blocks(e->n), nwg, k, e->wall_hist.p);
  int halted = 0;
  UpdState st;
  d2h(e, &halted, e->halted.p, 1);
  d2h(e, &st, e->upd.p, 1);
  std::vector<double> h((size_t)k * nwg * 3);
  if (nwg) d2h(e, h.data(), e->wall_hist.p, h.size());
  sync(e);
  const int done = halted ? halted - 1 : k;
  *reason = halted ? st.halt_reason : 0;
  collect_timing(e, done);
  if (done < k) {  // walls stay where the last completed step left them
    const std::vector<double>& o = done > 0 ? wall_o[done - 1] : before;
    for (size_t w = 0; w < e->walls.size(); w++)
      for (int c = 0; c < 3; c++) e->walls[w].o[c] = o[3 * w + c];
  }
  if (nwg) record_walls(e, h, done);
  if (done & 1) swap_history_buffers(e);
  e->t += done;
  e->pending_host = done < k;
  return done;
}

void run_steps(esp_engine* e, int nsteps)
{
  if (!e->started) {
    refresh_orders(e);
    if (e->wg.size() > ESP_MAX_WALL_GROUPS)
      throw EspError(ESP_ERR_CAPACITY, "more than 8 wall interaction groups");
    if (e->grav.size() > ESP_MAX_GRAVITY)
      throw EspError(ESP_ERR_CAPACITY, "more than 4 gravity groups");
    if (e->bg.size() > ESP_MAX_BOND_GROUPS)
      throw EspError(ESP_ERR_CAPACITY, "more than 8 bonded groups");
    if (e->bg.size() + e->dg.size() > ESP_MAX_FORCE_SEGS)
      throw EspError(ESP_ERR_CAPACITY, "more than 16 interaction groups");
    for (auto& G : e->dg)
      if (G.ex.size() > ESP_MAX_EXCL)
        throw EspError(ESP_ERR_CAPACITY, "more than 4 exclusions on one group");
    upload(e);
  }
  int done = 0;
  while (done < nsteps) {
    if (e->first_time || e->pending_host) {
      host_step(e, e->first_time ? 0 : e->pending_reason);
      e->pending_host = false;
      done++;
      continue;
    }
    // queue depth adapts to how often host bookkeeping interrupts the queue:
    // it doubles after a full batch and drops to twice the completed run
    const int k = std::min(e->kcur, nsteps - done);
    int reason = 0;
    const int got = fast_batch(e, k, &reason);
    e->pending_reason = reason;
    e->kcur = got == k ? std::min(e->K, 2 * e->kcur) : std::max(1, std::min(e->K, 2 * (got + 1)));
    done += got;
  }
}

// ---------------------------------------------------------------------------
// queries
// ---------------------------------------------------------------------------
std::vector<int> id_order(esp_engine* e)
{
  std::vector<int> order(e->n);
  for (int i = 0; i < e->n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return e->h_id[a] < e->h_id[b]; });
  return order;
}

// Read-only views for queries.  Dead bonds stay in storage and appended pairs
// stay in their appendix until host bookkeeping folds them (fold), so queries
// assemble the lists the way fold would and leave the device state alone.
std::vector<char> bond_alive(esp_engine* e, BondGroup& G)
{
  std::vector<char> a(G.n, 1);
  if (!e->started || G.n == 0) return a;
  std::vector<int2> ij(G.n);
  d2h(e, ij.data(), G.ij.p, G.n);
  sync(e);
  for (int b = 0; b < G.n; b++) a[b] = ij[b].x >= 0;
  return a;
}

int bond_live_count(esp_engine* e, BondGroup& G)
{
  int c = 0;
  for (char a : bond_alive(e, G)) c += a;
  return c;
}

int dyn_appended(esp_engine* e, DynGroup& G)
{
  if (!e->started) return 0;
  int na = 0;
  d2h(e, &na, G.napp.p, 1);
  sync(e);
  return na;
}

// slot-pair key -> id-pair key
long long id_key(esp_engine* e, long long skey)
{
  return pkey(e->h_id[(int)(skey >> 32)], e->h_id[(int)(skey & 0xffffffffLL)]);
}

// stored pairs of a dynamic group (main list and appendix) as id keys in
// ascending order; equal keys (NRotElastic may store a pair twice) keep
// storage order
void dyn_view(esp_engine* e, DynGroup& G, std::vector<long long>& key, std::vector<double4>* ff)
{
  const int tot = G.n + dyn_appended(e, G);
  key.assign(tot, 0);
  std::vector<long long> k(std::max(tot, 1));
  std::vector<double4> f(ff ? std::max(tot, 1) : 0);
  if (tot > 0) {
    d2h(e, k.data(), G.key.p, tot);
    if (ff) d2h(e, f.data(), G.ff.p, tot);
    sync(e);
  }
  for (int i = 0; i < tot; i++) k[i] = id_key(e, k[i]);
  std::vector<int> perm(tot);
  for (int i = 0; i < tot; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return k[a] < k[b]; });
  if (ff) ff->resize(tot);
  for (int i = 0; i < tot; i++) {
    key[i] = k[perm[i]];
    if (ff) (*ff)[i] = f[perm[i]];
  }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
#define API_BEGIN try {
#define API_END                                                  \
  }                                                              \
  catch (const EspError& ex)                                     \
  {                                                              \
    e->err = ex.what();                                          \
    return ex.code;                                              \
  }                                                              \
  catch (const std::exception& ex)                               \
  {                                                              \
    e->err = ex.what();                                          \
    return ESP_ERR_STATE;                                        \
  }

namespace {
BondGroup& add_bond_group(esp_engine* e, const char* name, const esp_rot_bond_prms* prm,
                          const double* per_bond_max, int n_per_bond)
{
  require_setup(e, "createInteractionGroup");
  if (!e->rot) throw EspError(ESP_ERR_STATE, "rotational bonds need particle type RotSphere");
  if (!name || !prm) throw EspError(ESP_ERR_ARG, "bad group");
  if (name_taken(e, name)) throw EspError(ESP_ERR_ARG, std::string("interaction group '") + name + "' exists");
  BondGroup G;
  G.name = name;
  G.prm = *prm;
  // connections of this tag in arrival order; repeated id pairs are rejected by
  // isIn at their second occurrence (pi_storage_e.hpp:216-250)
  std::vector<int> cidx;
  for (size_t c = 0; c < e->h_conn.size(); c += 3)
    if (e->h_conn[c] == prm->tag) cidx.push_back((int)c);
  if (per_bond_max && (int)cidx.size() != n_per_bond)
    throw EspError(ESP_ERR_ARG, "per_bond_max has " + std::to_string(n_per_bond) + " entries for " +
                                    std::to_string(cidx.size()) + " connections");
  std::unordered_map<long long, int> seen;
  for (size_t j = 0; j < cidx.size(); j++) {
    const int a = e->h_conn[cidx[j] + 1], b = e->h_conn[cidx[j] + 2];
    const int id1 = std::min(a, b), id2 = std::max(a, b);
    const long long key = pkey(id1, id2);
    if (seen.count(key)) continue;
    seen[key] = 1;
    auto i1 = e->id2slot.find(id1), i2 = e->id2slot.find(id2);
    if (i1 == e->id2slot.end() || i2 == e->id2slot.end()) continue;
    const int s1 = i1->second, s2 = i2->second;
    double maxes[4] = {prm->max_n, prm->max_s, prm->max_t, prm->max_b};
    if (per_bond_max) memcpy(maxes, per_bond_max + 4 * j, sizeof maxes);
    // circular x: the initial bond vector runs to the image of particle 2
    double x2 = e->h_pos[3 * s2];
    if (e->grid.circ_x) {
      const double d = e->h_pos[3 * s1] - x2, L = e->grid.period;
      if (d > 0.5 * L) x2 = x2 + L;
      else if (d < -0.5 * L) x2 = x2 - L;
    }
    const BondConst bc = esp_bond_constants(
        mk3(e->h_pos[3 * s1], e->h_pos[3 * s1 + 1], e->h_pos[3 * s1 + 2]), e->h_rad[s1],
        mk3(x2, e->h_pos[3 * s2 + 1], e->h_pos[3 * s2 + 2]), e->h_rad[s2], prm->kr,
        prm->ks, prm->kt, prm->kb, maxes, prm->scaling, prm->meanR_scaling, e->do2d);
    G.hid1.push_back(id1);
    G.hid2.push_back(id2);
    G.h_k4.push_back(make_double4(bc.kr, bc.ks, bc.kt, bc.kb));
    G.h_mx4.push_back(make_double4(bc.max_n, bc.max_s, bc.max_t, bc.max_b));
    G.h_D4.push_back(make_double4(bc.D.x, bc.D.y, bc.D.z, 0.0));
  }
  G.n = (int)G.hid1.size();
  e->bg.push_back(std::move(G));
  return e->bg.back();
}
}  // namespace


// ===========================================================================
// restart checkpoints in the reference's text format
// ===========================================================================
namespace {
// Foundation/version.h:19-20, bzrversion.h:1
const int kCheckPointVersion = 3;
const char* kVersionLine = "ESyS-Particle_2.3.5_bzr_rev. 1218";

std::string cp_file_name(const std::string& prefix, long long t, int rank)
{
  std::ostringstream s;
  s << prefix << "_t=" << t << "_" << rank << ".txt";  // CheckPointParams.cpp:37-43
  return s.str();
}

// TSubLattice::saveCheckPointData (SubLattice.hpp:3599-3672) for the one subdomain
void cp_write(esp_engine* e, const std::string& prefix, int num_time_steps, int precision)
{
  if (!e->started) refresh_orders(e);
  const int n = e->n;
  std::vector<double4> P(n), V(n), O(n), F(n), A(n), Q(n);
  std::vector<unsigned> fl(n);
  if (e->started) {
    d2h(e, P.data(), e->pos.p, n);
    d2h(e, V.data(), e->vel.p, n);
    d2h(e, O.data(), e->old.p, n);
    d2h(e, F.data(), e->frc.p, n);
    if (e->rot) {
      d2h(e, A.data(), e->ang.p, n);
      d2h(e, Q.data(), e->quat.p, n);
    }
    d2h(e, fl.data(), e->flags.p, n);
    sync(e);
  } else {
    for (int i = 0; i < n; i++) {
      P[i] = make_double4(e->h_pos[3 * i], e->h_pos[3 * i + 1], e->h_pos[3 * i + 2], e->h_rad[i]);
      V[i] = make_double4(e->h_vel[3 * i], e->h_vel[3 * i + 1], e->h_vel[3 * i + 2], 0.0);
      O[i] = make_double4(P[i].x, P[i].y, P[i].z, e->h_mass[i]);
      F[i] = make_double4(0.0, 0.0, 0.0, 0.0);
      A[i] = make_double4(e->h_ang[3 * i], e->h_ang[3 * i + 1], e->h_ang[3 * i + 2], 0.0);
      Q[i] = make_double4(e->h_quat[4 * i], e->h_quat[4 * i + 1], e->h_quat[4 * i + 2],
                          e->h_quat[4 * i + 3]);
      fl[i] = e->h_flags[i];
    }
  }
  const std::string data_name = cp_file_name(prefix, e->t, 1);
  {
    // CheckPointInfo::write (CheckPointInfo.cpp:73-96), default stream precision
    std::ofstream o(cp_file_name(prefix, e->t, 0));
    if (!o) throw EspError(ESP_ERR_ARG, "cannot write " + cp_file_name(prefix, e->t, 0));
    o << "V " << kCheckPointVersion << "\n" << num_time_steps << "\n" << e->dt << "\n" << e->t
      << "\n" << kVersionLine << "\n";
    // GeometryInfo::writeWithoutVersion (GeometryInfo.cpp:134-155)
    o << "BoundingBox " << e->cfg.domain_min[0] << " " << e->cfg.domain_min[1] << " "
      << e->cfg.domain_min[2] << " " << e->cfg.domain_max[0] << " " << e->cfg.domain_max[1]
      << " " << e->cfg.domain_max[2] << "\n";
    o << "PeriodicBoundaries " << (e->grid.circ_x ? 1 : 0) << " 0 0\nDimension " << (e->do2d ? "2D" : "3D") << "\n" << data_name;
  }
  std::ofstream o(data_name);
  if (!o) throw EspError(ESP_ERR_ARG, "cannot write " + data_name);
  o.precision(precision);
  // ParallelParticleArray::saveCheckPointData (pp_array.hpp:1216-1235)
  o << e->grid.x0 << " " << e->grid.y0 << " " << e->grid.z0 << "\n";
  o << e->grid.gx << " " << e->grid.gy << " " << e->grid.gz << "\n";
  o << n << "\n";
  for (int i : id_order(e)) {
    // CParticle / CRotParticle::saveCheckPointData (Particle.cpp:461-475, RotParticle.cpp:510-519)
    const double m = O[i].w, r = P[i].w;
    o << P[i].x << " " << P[i].y << " " << P[i].z << " " << r << " " << e->h_id[i] << " "
      << e->h_tag[i] << " " << m << " " << e->h_init[3 * i] << " " << e->h_init[3 * i + 1] << " "
      << e->h_init[3 * i + 2] << " " << O[i].x << " " << O[i].y << " " << O[i].z << " " << V[i].x
      << " " << V[i].y << " " << V[i].z << " " << F[i].x << " " << F[i].y << " " << F[i].z
      << " 0 0 0 " << ((fl[i] & PF_DYN) ? 1 : 0);
    if (e->rot) {
      const double I = e->h_inert[i] != 0.0 ? e->h_inert[i]
                                            : (e->do2d ? 0.5 * m * r * r : 0.4 * m * r * r);
      o << " " << I << " " << Q[i].x << " " << Q[i].y << " " << Q[i].z << " " << Q[i].w << " "
        << A[i].x << " " << A[i].y << " " << A[i].z << " " << ((fl[i] & PF_ROT) ? 1 : 0);
    }
    o << "\n";
  }
  // bonded groups, name order (PIS_EB::saveCheckPointData, pi_storage_eb.hpp:78-97)
  o << e->bg.size() << "\n";
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    std::vector<int2> ij(G.n);
    std::vector<int> rank(G.n);
    std::vector<double4> k4(G.n), mx4(G.n), D4(G.n);
    if (e->started && G.n > 0) {
      d2h(e, ij.data(), G.ij.p, G.n);
      d2h(e, rank.data(), G.rank.p, G.n);
      d2h(e, k4.data(), G.k4.p, G.n);
      d2h(e, mx4.data(), G.mx4.p, G.n);
      d2h(e, D4.data(), G.D4.p, G.n);
      sync(e);
    } else {
      for (int b = 0; b < G.n; b++) {
        ij[b] = make_int2(0, 0);
        rank[b] = b;
        k4[b] = G.h_k4[b];
        mx4[b] = G.h_mx4[b];
        D4[b] = G.h_D4[b];
      }
    }
    std::vector<int> live;
    for (int b = 0; b < G.n; b++)
      if (ij[b].x >= 0) live.push_back(b);
    std::sort(live.begin(), live.end(), [&](int a, int b) { return rank[a] < rank[b]; });
    o << "RotBonded\n" << live.size();  // BrittleBeamSC's type string is RotBonded too
    for (int b : live) {
      o << "\n" << G.hid1[b] << " " << G.hid2[b] << " 1 " << k4[b].x << " " << k4[b].y << " "
        << k4[b].w << " " << k4[b].z << " ";
      if (G.sc) {
        // BrittleBeamSCInteraction::saveRestartData (BrittleBeamSC.cpp:287-302)
        o << G.prm.scaling << " " << D4[b].x << " " << D4[b].y << " " << D4[b].z << " "
          << G.prm.tag << " " << G.sc_sigma3 << " " << G.sc_cohesion << " " << G.sc_tan
          << std::endl;
      } else {
        // CRotBondedInteraction::saveRestartData (RotBondedInteraction.cpp:379-399)
        o << G.prm.scaling << " " << G.prm.meanR_scaling << " " << G.prm.truncated << " "
          << G.prm.beta1 << " " << G.prm.beta2 << " " << mx4[b].x << " " << mx4[b].y << " "
          << mx4[b].z << " " << mx4[b].w << " " << D4[b].x << " " << D4[b].y << " " << D4[b].z
          << " " << G.prm.tag << std::endl;
      }
    }
    o << "\n";
  }
  // dynamic groups that save (PIS_ED: RotFriction), name order
  // (PIS_ED::saveCheckPointData, pi_storage_ed.hpp:201-218).  The reference writes
  // the first record straight after the count; a newline separates them here.
  int nsave = 0;
  for (auto& G : e->dg) nsave += G.kind == 3;
  o << nsave << "\n";
  for (int gi : e->dord) {
    DynGroup& G = e->dg[gi];
    if (G.kind != 3) continue;
    std::vector<long long> keys;
    std::vector<double4> ff;
    if (e->started) {
      dyn_view(e, G, keys, &ff);
    } else {
      for (size_t k = 0; k < G.pre_id1.size(); k++) {
        keys.push_back(pkey(G.pre_id1[k], G.pre_id2[k]));
        ff.push_back(G.pre_ff[k]);
      }
    }
    o << "RotFriction" << std::endl << keys.size() << "\n";
    for (size_t k = 0; k < keys.size(); k++) {
      const int id1 = (int)(keys[k] >> 32), id2 = (int)(keys[k] & 0xffffffffLL);
      const int s1 = e->id2slot[id1], s2 = e->id2slot[id2];
      const double r1 = e->h_rad[s1], r2 = e->h_rad[s2];
      double kk, kks;
      esp_friction_k(G.k, G.ks, G.scaling, G.meanR, r1, r2, e->do2d, kk, kks);
      const int flg = (int)ff[k].w;
      // CRotFrictionInteraction::saveRestartData (RotFricInteraction.cpp:597-614)
      o << kk << " " << (r1 + r2) << " " << G.mu_d << " " << G.mu_s << " " << kks << " " << e->dt
        << " " << G.scaling << " " << G.meanR << " " << id1 << " " << id2 << " " << (flg & 1)
        << " " << ((flg >> 1) & 1) << " " << ff[k].x << " " << ff[k].y << " " << ff[k].z
        << std::endl;
    }
    o << "\n";
  }
  // walls, name order (SubLattice.hpp:3636-3643, CWall::writeCheckPoint, Wall.cpp:118-121)
  std::vector<int> word(e->walls.size());
  for (size_t w = 0; w < word.size(); w++) word[w] = (int)w;
  std::sort(word.begin(), word.end(),
            [&](int a, int b) { return e->walls[a].name < e->walls[b].name; });
  o << "Walls " << e->walls.size() << "\n";
  for (int w : word) {
    const Wall& W = e->walls[w];
    o << W.name << "\n"
      << W.o[0] << " " << W.o[1] << " " << W.o[2] << " " << W.old[0] << " " << W.old[1] << " "
      << W.old[2] << " " << W.n[0] << " " << W.n[1] << " " << W.n[2] << "\n";
  }
  o << "Spheres 0\nTriMesh 0\nMesh2D 0\n";
  if (!o) throw EspError(ESP_ERR_ARG, "writing " + data_name + " failed");
}

// whitespace-separated tokens of a file
std::vector<std::string> cp_tokens(const std::string& path)
{
  std::ifstream f(path);
  if (!f) throw EspError(ESP_ERR_ARG, "cannot open " + path);
  std::vector<std::string> t;
  std::string s;
  while (f >> s) t.push_back(s);
  return t;
}

struct CpReader {
  std::vector<std::string> t;
  size_t k = 0;
  std::string file;
  const std::string& next()
  {
    if (k >= t.size()) throw EspError(ESP_ERR_ARG, file + ": unexpected end of checkpoint data");
    return t[k++];
  }
  double d() { return std::stod(next()); }
  long long i() { return std::stoll(next()); }
  void expect(const char* w)
  {
    const std::string& s = next();
    if (s != w) throw EspError(ESP_ERR_ARG, file + ": expected '" + w + "', found '" + s + "'");
  }
};

// CLatticeMaster::loadCheckPointData (LatticeMaster.cpp:780-820) and
// TSubLattice::loadCheckPointData (SubLattice.hpp:3675-3790)
void cp_load(esp_engine* e, const std::string& info_path)
{
  refresh_orders(e);
  CpReader in{cp_tokens(info_path), 0, info_path};
  in.expect("V");
  in.i();  // checkpoint format version
  in.i();  // number of time steps of the saving run
  in.d();  // time step size of the saving run (the script sets its own)
  const long long t0 = in.i();
  const std::string head = in.next();
  if (head.compare(0, 4, "ESyS") != 0 && head != "LSMGeometry")
    throw EspError(ESP_ERR_ARG, info_path + ": not a checkpoint info file");
  in.next();  // version
  in.expect("BoundingBox");
  for (int c = 0; c < 6; c++) in.d();
  in.expect("PeriodicBoundaries");
  for (int c = 0; c < 3; c++)
    if (in.i() != (c == 0 ? e->grid.circ_x : 0))
      throw EspError(ESP_ERR_ARG, "the checkpoint's periodic dimensions differ from the domain's");
  in.expect("Dimension");
  in.next();
  std::vector<std::string> files(in.t.begin() + in.k, in.t.end());
  if (files.empty()) throw EspError(ESP_ERR_ARG, info_path + ": no subdomain files listed");
  const size_t slash = info_path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "" : info_path.substr(0, slash + 1);
  for (const std::string& name : files) {
    std::string path = name;
    if (!std::ifstream(path) && !dir.empty()) path = dir + name.substr(name.find_last_of('/') + 1);
    CpReader r{cp_tokens(path), 0, path};
    for (int c = 0; c < 6; c++) r.d();  // table base point and base index
    const long long np = r.i();
    for (long long q = 0; q < np; q++) {
      double pos[3], vel[3], init[3], v3[3];
      for (int c = 0; c < 3; c++) pos[c] = r.d();
      const double rad = r.d();
      const int id = (int)r.i();
      const int tag = (int)r.i();
      const double mass = r.d();
      for (int c = 0; c < 3; c++) init[c] = r.d();
      for (int c = 0; c < 3; c++) v3[c] = r.d();  // position at the last search
      for (int c = 0; c < 3; c++) vel[c] = r.d();
      for (int c = 0; c < 3; c++) v3[c] = r.d();  // force of the last step
      for (int c = 0; c < 3; c++) v3[c] = r.d();  // circular shift
      const bool dyn = r.i() != 0;
      if (esp_add_particles(e, 1, &id, &tag, pos, &rad, &mass, vel) != ESP_OK)
        throw EspError(ESP_ERR_ARG, e->err);
      auto it = e->id2slot.find(id);
      const int s = it == e->id2slot.end() ? -1 : it->second;  // -1: outside the table
      if (s >= 0)
        for (int c = 0; c < 3; c++) e->h_init[3 * s + c] = init[c];
      if (e->rot) {
        double rs[8];
        for (int c = 0; c < 8; c++) rs[c] = r.d();  // inertia, quaternion, angular velocity
        const bool rot = r.i() != 0;
        if (s < 0) continue;
        e->h_inert[s] = rs[0];
        for (int c = 0; c < 4; c++) e->h_quat[4 * s + c] = rs[1 + c];
        for (int c = 0; c < 3; c++) e->h_ang[3 * s + c] = rs[5 + c];
        if (!rot) e->h_flags[s] &= ~PF_ROT;
      }
      if (s >= 0 && !dyn) e->h_flags[s] &= ~(PF_DYN | PF_ROT);
    }
    // bonded groups, matched to the script's groups in name order
    const long long nbg = r.i();
    if (nbg != (long long)e->bg.size())
      throw EspError(ESP_ERR_ARG, "the checkpoint holds " + std::to_string(nbg) +
                                      " bonded groups, the script created " +
                                      std::to_string(e->bg.size()));
    for (int gi : e->bord) {
      BondGroup& G = e->bg[gi];
      r.expect("RotBonded");
      const long long nb = r.i();
      std::unordered_map<long long, int> have;
      for (int b = 0; b < (int)G.hid1.size(); b++) have[pkey(G.hid1[b], G.hid2[b])] = 1;
      for (long long b = 0; b < nb; b++) {
        int id1 = (int)r.i(), id2 = (int)r.i();
        r.i();  // initialised flag
        const double kr = r.d(), ks = r.d(), kb = r.d(), kt = r.d();
        double mx[4] = {1.0, 1.0, 1.0, 1.0}, D[3];
        if (G.sc) {
          r.i();  // scaling
          for (int c = 0; c < 3; c++) D[c] = r.d();
          r.i();  // tag
          G.sc_sigma3 = r.d();
          G.sc_cohesion = r.d();
          G.sc_tan = r.d();
        } else {
          r.i();  // scaling
          r.i();  // mean-radius scaling
          for (int c = 0; c < 3; c++) r.d();  // truncation, beta1, beta2 (the group's apply)
          for (int c = 0; c < 4; c++) mx[c] = r.d();
          for (int c = 0; c < 3; c++) D[c] = r.d();
          r.i();  // tag
        }
        if (id1 > id2) std::swap(id1, id2);
        if (!e->id2slot.count(id1) || !e->id2slot.count(id2)) continue;
        if (!have.emplace(pkey(id1, id2), 1).second) continue;  // isIn (pi_storage_e.hpp:216-250)
        G.hid1.push_back(id1);
        G.hid2.push_back(id2);
        G.h_k4.push_back(make_double4(kr, ks, kt, kb));
        G.h_mx4.push_back(make_double4(mx[0], mx[1], mx[2], mx[3]));
        G.h_D4.push_back(make_double4(D[0], D[1], D[2], 0.0));
      }
      G.n = (int)G.hid1.size();
    }
    // dynamic groups that save: the script's RotFriction groups in name order
    const long long ndg = r.i();
    std::vector<int> fric;
    for (int gi : e->dord)
      if (e->dg[gi].kind == 3) fric.push_back(gi);
    if (ndg != (long long)fric.size())
      throw EspError(ESP_ERR_ARG, "the checkpoint holds " + std::to_string(ndg) +
                                      " saved dynamic groups, the script created " +
                                      std::to_string(fric.size()) + " RotFriction groups");
    for (int gi : fric) {
      DynGroup& G = e->dg[gi];
      r.expect("RotFriction");
      // the reference's writer puts the first record right after the count:
      // the section's records are 15 fields each up to the next section name
      size_t end = r.k;
      while (end < r.t.size() && r.t[end] != "RotFriction" && r.t[end] != "Walls") end++;
      const size_t nt = end - r.k;
      long long nc = 0;
      if (nt % 15 == 1) {
        nc = r.i();
      } else if (nt % 15 == 0 && nt > 0) {
        nc = (long long)(nt / 15);
        const std::string c = std::to_string(nc);
        if (r.t[r.k].compare(0, c.size(), c) != 0)
          throw EspError(ESP_ERR_ARG, r.file + ": RotFriction count and records disagree");
        r.t[r.k] = r.t[r.k].substr(c.size());
      } else {
        throw EspError(ESP_ERR_ARG, r.file + ": malformed RotFriction section");
      }
      for (long long c = 0; c < nc; c++) {
        for (int q = 0; q < 8; q++) r.d();  // k, r0, mu_d, mu_s, ks, dt, scaling, mean-R scaling
        int id1 = (int)r.i(), id2 = (int)r.i();
        const int slip = (int)r.i(), touch = (int)r.i();
        const double fx = r.d(), fy = r.d(), fz = r.d();
        if (id1 > id2) std::swap(id1, id2);
        G.pre_id1.push_back(id1);
        G.pre_id2.push_back(id2);
        G.pre_ff.push_back(make_double4(fx, fy, fz, (double)((slip ? 1 : 0) | (touch ? 2 : 0))));
      }
    }
    // walls
    r.expect("Walls");
    const long long nw = r.i();
    for (long long w = 0; w < nw; w++) {
      const std::string name = r.next();
      double v[9];
      for (int c = 0; c < 9; c++) v[c] = r.d();
      int wi = find_wall(e, name.c_str());
      if (wi < 0) {
        Wall W;
        W.name = name;
        for (int c = 0; c < 3; c++) W.step[c] = 0.0;
        e->walls.push_back(W);
        wi = (int)e->walls.size() - 1;
      }
      Wall& W = e->walls[wi];
      for (int c = 0; c < 3; c++) {
        W.o[c] = v[c];
        W.old[c] = v[3 + c];
        W.n[c] = v[6 + c];
      }
    }
    for (const char* kind : {"Spheres", "TriMesh", "Mesh2D"}) {
      r.expect(kind);
      if (r.i() != 0) throw EspError(ESP_ERR_ARG, std::string(kind) + " are not supported");
    }
  }
  e->t = t0;
}
}  // namespace

extern "C" {

const char* esp_create_error(void) { return g_create_err.c_str(); }

const char* esp_last_error(const esp_engine* e) { return e ? e->err.c_str() : g_create_err.c_str(); }

int esp_create(const esp_config* cfg, esp_engine** out)
{
  g_create_err.clear();
  if (!cfg || !out) {
    g_create_err = "null argument";
    return ESP_ERR_ARG;
  }
  if (cfg->particle_type != ESP_PTYPE_NROT && cfg->particle_type != ESP_PTYPE_ROT) {
    g_create_err = "unknown particle type";
    return ESP_ERR_ARG;
  }
  if (!(cfg->grid_spacing > 0) || !(cfg->verlet_dist >= 0)) {
    g_create_err = "gridSpacing must be > 0 and verletDist >= 0";
    return ESP_ERR_ARG;
  }
  // TSubLattice::initNeighborTable (SubLattice.hpp:455-478) and the 1x1x1
  // ParallelParticleArray table (pp_array.hpp:61-103)
  double nr = cfg->grid_spacing;
  double mn[3], mx[3], nmin[3], nmax[3];
  for (int c = 0; c < 3; c++) { mn[c] = cfg->domain_min[c]; mx[c] = cfg->domain_max[c]; }
  double sz[3];
  if (cfg->circular_x) {
    // TSubLattice::initNeighborTable(min, max, circ) (SubLattice.hpp:502-515): the
    // range is adjusted so that whole cells span the circular x size
    const double xs = mx[0] - mn[0];
    if (std::fabs(xs / nr - (double)lrint(xs / nr)) > 1e-6) nr = xs / std::floor(xs / nr);
  }
  for (int c = 0; c < 3; c++) {
    sz[c] = mx[c] - mn[c];
    if (!(c == 0 && cfg->circular_x)) sz[c] = nr * std::ceil(sz[c] / nr);
  }
  for (int c = 0; c < 3; c++) {
    const double grow = sz[c] - (mx[c] - mn[c]);
    nmin[c] = mn[c] - grow * 0.5;
    nmax[c] = mx[c] + grow * 0.5;
  }
  long ng[3];
  for (int c = 0; c < 3; c++) {
    ng[c] = lrint((nmax[c] - nmin[c]) / nr);
    if (std::fabs((double)ng[c] - (nmax[c] - nmin[c]) / nr) > 1e-6 || ng[c] < 1) {
      g_create_err = "size doesn't fit range";
      return ESP_ERR_ARG;
    }
  }
  const double ncells = (double)(ng[0] + 2) * (double)(ng[1] + 2) * (double)(ng[2] + 2);
  if (ncells > 2.0e8) {
    g_create_err = "neighbour table too large";
    return ESP_ERR_CAPACITY;
  }
  if (cfg->circular_x && ng[0] < 3) {
    g_create_err = "a circular x dimension needs at least 3 cells";
    return ESP_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_create_err = "no HIP device available";
    return ESP_ERR_DEVICE;
  }
  if (cfg->device < 0 || cfg->device >= ndev) {
    g_create_err = "bad device ordinal";
    return ESP_ERR_ARG;
  }
  if (hipSetDevice(cfg->device) != hipSuccess) {
    g_create_err = "hipSetDevice failed";
    return ESP_ERR_DEVICE;
  }
  esp_engine* e = new esp_engine();
  e->cfg = *cfg;
  e->rot = cfg->particle_type == ESP_PTYPE_ROT;
  e->do2d = cfg->do_2d;
  e->dt = cfg->dt;
  e->alpha = cfg->verlet_dist;
  e->K = std::max(1, std::min(cfg->steps_per_batch > 0 ? cfg->steps_per_batch : 32, 4096));
  e->grid.x0 = nmin[0]; e->grid.y0 = nmin[1]; e->grid.z0 = nmin[2];
  e->grid.dim = nr;
  e->grid.gx = e->grid.gy = e->grid.gz = -1;
  e->grid.nx = (int)ng[0] + 2; e->grid.ny = (int)ng[1] + 2; e->grid.nz = (int)ng[2] + 2;
  e->grid.circ_x = 0;
  e->grid.period = 0.0;
  if (cfg->circular_x) {
    // one table over the whole circular x range: no padding cells, the
    // stencil wraps, interactions see the nearest image
    e->grid.circ_x = 1;
    e->grid.gx = 0;
    e->grid.nx = (int)ng[0];
    e->grid.period = nmax[0] - nmin[0];
  }
  e->ncell = e->grid.nx * e->grid.ny * e->grid.nz;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    g_create_err = "hipStreamCreate failed";
    delete e;
    return ESP_ERR_DEVICE;
  }
  // the allocation-free servo's staging address on this device, looked up
  // here with the device current (a failure is reported by that entry point)
  try {
    (void)scratch_on(cfg->device);
  } catch (const EspError&) {
  }
  *out = e;
  return ESP_OK;
}

void esp_destroy(esp_engine* e)
{
  if (!e) return;
  if (e->stream) {
    (void)hipStreamSynchronize(e->stream);
  }
  for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
  hipStream_t s = e->stream;
  delete e;
  if (s) (void)hipStreamDestroy(s);
}

int esp_set_time_step_size(esp_engine* e, double dt)
{
  e->dt = dt;
  return ESP_OK;
}

int esp_add_particles(esp_engine* e, int n, const int* id, const int* tag, const double* pos,
                      const double* rad, const double* mass, const double* vel)
{
  API_BEGIN
  require_setup(e, "createParticles");
  if (n < 0 || (n > 0 && (!id || !pos || !rad))) throw EspError(ESP_ERR_ARG, "bad particle arrays");
  std::unordered_map<int, int> added;
  for (int i = 0; i < n; i++) {
    if (id[i] < 0) throw EspError(ESP_ERR_ARG, "particle ids must be >= 0");
    if (e->id2slot.count(id[i]) || added.count(id[i]))
      throw EspError(ESP_ERR_ARG, "duplicate particle id " + std::to_string(id[i]));
    added[id[i]] = 1;
  }
  for (int i = 0; i < n; i++) {
    // NeighborTable::insert (ntable.hpp:231-255): particles outside are dropped
    if (host_cell(e, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]) < 0) continue;
    const double r = rad[i];
    const int slot = e->n++;
    e->id2slot[id[i]] = slot;
    e->h_id.push_back(id[i]);
    e->h_tag.push_back(tag ? tag[i] : 0);
    for (int c = 0; c < 3; c++) e->h_pos.push_back(pos[3 * i + c]);
    for (int c = 0; c < 3; c++) e->h_vel.push_back(vel ? vel[3 * i + c] : 0.0);
    e->h_rad.push_back(r);
    e->h_mass.push_back(mass ? mass[i] : (e->do2d ? r * r : r * r * r));
    e->h_flags.push_back(PF_DYN | (e->rot ? PF_ROT : 0u));
    for (int c = 0; c < 3; c++) e->h_init.push_back(pos[3 * i + c]);
    e->h_quat.insert(e->h_quat.end(), {1.0, 0.0, 0.0, 0.0});
    e->h_ang.insert(e->h_ang.end(), {0.0, 0.0, 0.0});
    e->h_inert.push_back(0.0);
    e->h_order.push_back(e->next_order++);
    e->ord_valid = false;
  }
  return ESP_OK;
  API_END
}

int esp_set_particle_density(esp_engine* e, int tag, int mask, double rho)
{
  API_BEGIN
  require_setup(e, "setParticleDensity");
  const double pi = 3.141592654;
  for (int i = 0; i < e->n; i++) {
    if ((e->h_tag[i] & mask) != (tag & mask)) continue;
    const double r = e->h_rad[i];
    e->h_mass[i] = e->do2d ? rho * pi * r * r : (4.0 / 3.0) * rho * pi * r * r * r;
  }
  return ESP_OK;
  API_END
}

int esp_set_particle_non_dynamic(esp_engine* e, int tag)
{
  API_BEGIN
  require_setup(e, "setParticleNonDynamic");
  for (int i = 0; i < e->n; i++)
    if (e->h_tag[i] == tag) e->h_flags[i] &= ~(PF_DYN | PF_ROT);
  return ESP_OK;
  API_END
}

int esp_set_particle_non_rotational(esp_engine* e, int tag)
{
  API_BEGIN
  require_setup(e, "setParticleNonRotational");
  for (int i = 0; i < e->n; i++)
    if (e->h_tag[i] == tag) e->h_flags[i] &= ~PF_ROT;
  return ESP_OK;
  API_END
}

namespace {
// particle id -> slot (arrival order before the first step, spatial order after)
int slot_of_id(esp_engine* e, int id)
{
  if (e->started) {
    auto it = e->id2slot.find(id);
    return it == e->id2slot.end() ? -1 : it->second;
  }
  for (int i = 0; i < e->n; i++)
    if (e->h_id[i] == id) return i;
  return -1;
}

// writes the xyz part of one particle's double4 field on the device
void set_xyz(esp_engine* e, double4* field, int slot, const double v[3])
{
  HIPCHK(hipMemcpyAsync(reinterpret_cast<double*>(field + slot), v, 3 * sizeof(double),
                        hipMemcpyHostToDevice, e->stream));
  sync(e);
}
}  // namespace

int esp_set_particle_vel(esp_engine* e, int id, const double v[3])
{
  API_BEGIN
  if (!v) throw EspError(ESP_ERR_ARG, "no velocity");
  const int s = slot_of_id(e, id);
  if (s < 0) throw EspError(ESP_ERR_ARG, "unknown particle id " + std::to_string(id));
  if (e->started) set_xyz(e, e->vel.p, s, v);
  else
    for (int c = 0; c < 3; c++) e->h_vel[3 * s + c] = v[c];
  return ESP_OK;
  API_END
}

int esp_set_particle_ang_vel(esp_engine* e, int id, const double w[3])
{
  API_BEGIN
  if (!w) throw EspError(ESP_ERR_ARG, "no angular velocity");
  if (!e->rot) throw EspError(ESP_ERR_ARG, "setParticleAngularVelocity needs rotational particles");
  const int s = slot_of_id(e, id);
  if (s < 0) throw EspError(ESP_ERR_ARG, "unknown particle id " + std::to_string(id));
  if (e->started) set_xyz(e, e->ang.p, s, w);
  else
    for (int c = 0; c < 3; c++) e->h_ang[3 * s + c] = w[c];
  return ESP_OK;
  API_END
}

int esp_add_connections(esp_engine* e, int n, const int* id1, const int* id2, const int* tag)
{
  API_BEGIN
  require_setup(e, "createConnections");
  if (n < 0 || (n > 0 && (!id1 || !id2))) throw EspError(ESP_ERR_ARG, "bad connection arrays");
  for (int i = 0; i < n; i++) {
    e->h_conn.push_back(tag ? tag[i] : 0);
    e->h_conn.push_back(id1[i]);
    e->h_conn.push_back(id2[i]);
  }
  return ESP_OK;
  API_END
}


int esp_add_rot_bond_group(esp_engine* e, const char* name, const esp_rot_bond_prms* prm,
                           const double* per_bond_max, int n_per_bond)
{
  API_BEGIN
  add_bond_group(e, name, prm, per_bond_max, n_per_bond);
  return ESP_OK;
  API_END
}

int esp_add_brittle_beam_sc_group(esp_engine* e, const char* name,
                                  const esp_brittle_beam_sc_prms* sc)
{
  API_BEGIN
  if (!sc) throw EspError(ESP_ERR_ARG, "bad group");
  // BrittleBeamSCIGP -> ARotBondedIGP(kr, ks, kt, kb, tag, scaling=true)
  // (BrittleBeamSC.cpp:46-56); strengths do not apply
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
  BondGroup& G = add_bond_group(e, name, &prm, nullptr, 0);
  G.sc = 1;
  G.sc_sigma3 = sc->tension_cutoff;
  G.sc_cohesion = sc->cohesion;
  G.sc_tan = std::tan(sc->friction_angle / 57.2957795);  // BrittleBeamSC.cpp:78
  return ESP_OK;
  API_END
}

static int add_dyn(esp_engine* e, const char* name, int kind, double k, double ks, double mud,
                   double mus, int scaling, int meanR)
{
  API_BEGIN
  require_setup(e, "createInteractionGroup");
  if (!name) throw EspError(ESP_ERR_ARG, "bad group name");
  if (name_taken(e, name)) throw EspError(ESP_ERR_ARG, std::string("interaction group '") + name + "' exists");
  DynGroup G;
  G.name = name;
  G.kind = kind;
  G.k = k; G.ks = ks; G.mu_d = mud; G.mu_s = mus;
  G.scaling = scaling; G.meanR = meanR;
  e->dg.push_back(std::move(G));
  return ESP_OK;
  API_END
}

int esp_add_nrot_elastic_group(esp_engine* e, const char* name, double k, int scaling)
{
  return add_dyn(e, name, 1, k, 0, 0, 0, scaling, 1);
}

int esp_add_rot_elastic_group(esp_engine* e, const char* name, double kr, int scaling)
{
  if (!e->rot) {
    e->err = "RotElastic needs particle type RotSphere";
    return ESP_ERR_STATE;
  }
  return add_dyn(e, name, 2, kr, 0, 0, 0, scaling, 1);
}

int esp_add_rot_friction_group(esp_engine* e, const char* name, const esp_rot_friction_prms* p)
{
  if (!e->rot) {
    e->err = "RotFriction needs particle type RotSphere";
    return ESP_ERR_STATE;
  }
  if (!p) {
    e->err = "null parameters";
    return ESP_ERR_ARG;
  }
  const int rc = add_dyn(e, name, 3, p->k_n, p->k_s, p->mu_d, p->mu_s, p->scaling, p->meanR_scaling);
  if (rc == ESP_OK) e->dg.back().rigid = p->rigid != 0;
  return rc;
}

int esp_add_exclusion(esp_engine* e, const char* excluding, const char* excluded)
{
  API_BEGIN
  require_setup(e, "createExclusion");
  const int b = find_bg(e, excluding ? excluding : "");
  const int d = find_dg(e, excluded ? excluded : "");
  if (b < 0) throw EspError(ESP_ERR_ARG, std::string("excluding group '") + (excluding ? excluding : "") + "' is not a bonded group");
  if (d < 0) throw EspError(ESP_ERR_ARG, std::string("excluded group '") + (excluded ? excluded : "") + "' is not a dynamic group");
  if (e->dg[d].ex.size() >= ESP_MAX_EXCL) throw EspError(ESP_ERR_CAPACITY, "too many exclusions");
  e->dg[d].ex.push_back(b);  // dynamic groups are empty before the first step: nothing to purge
  return ESP_OK;
  API_END
}

int esp_add_wall(esp_engine* e, const char* name, const double o[3], const double n[3])
{
  API_BEGIN
  if (!name || find_wall(e, name) >= 0) throw EspError(ESP_ERR_ARG, "bad or duplicate wall name");
  require_setup(e, "createWall");
  Wall W;
  W.name = name;
  for (int c = 0; c < 3; c++) { W.o[c] = o[c]; W.n[c] = n[c]; W.step[c] = 0.0; W.old[c] = o[c]; }
  e->walls.push_back(W);
  return ESP_OK;
  API_END
}

int esp_add_elastic_wall_group(esp_engine* e, const char* name, const char* wall, double k)
{
  API_BEGIN
  require_setup(e, "createInteractionGroup");
  const int w = find_wall(e, wall ? wall : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall '") + (wall ? wall : "") + "'");
  if (!name || name_taken(e, name)) throw EspError(ESP_ERR_ARG, "bad or duplicate group name");
  if (e->wg.size() >= ESP_MAX_WALL_GROUPS) throw EspError(ESP_ERR_CAPACITY, "more than 8 wall groups");
  e->wg.push_back(WallGroup{name, w, k});
  return ESP_OK;
  API_END
}

int esp_add_bonded_wall_group(esp_engine* e, const char* name, const char* wall, double k,
                              int tag, int mask)
{
  API_BEGIN
  require_setup(e, "createInteractionGroup");
  const int w = find_wall(e, wall ? wall : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall '") + (wall ? wall : "") + "'");
  if (!name || name_taken(e, name)) throw EspError(ESP_ERR_ARG, "bad or duplicate group name");
  if (e->wg.size() >= ESP_MAX_WALL_GROUPS) throw EspError(ESP_ERR_CAPACITY, "more than 8 wall groups");
  WallGroup G{name, w, k};
  G.bonded = 1;
  G.tag = tag;
  G.mask = mask;
  e->wg.push_back(G);
  return ESP_OK;
  API_END
}

int esp_set_wall_group_tags(esp_engine* e, const char* name, int tag, int mask)
{
  API_BEGIN
  require_setup(e, "createInteractionGroupTagged");
  for (auto& G : e->wg)
    if (G.name == (name ? name : "")) {
      if (G.bonded) throw EspError(ESP_ERR_ARG, "a bonded wall group selects its tags itself");
      G.tagged = 1;
      G.tag = tag;
      G.mask = mask;
      return ESP_OK;
    }
  throw EspError(ESP_ERR_ARG, "not an elastic wall group");
  API_END
}

int esp_add_gravity(esp_engine* e, const char* name, const double g[3])
{
  API_BEGIN
  require_setup(e, "createInteractionGroup");
  if (!name || name_taken(e, name)) throw EspError(ESP_ERR_ARG, "bad or duplicate group name");
  if (e->grav.size() >= ESP_MAX_GRAVITY) throw EspError(ESP_ERR_CAPACITY, "more than 4 gravity groups");
  Gravity G;
  G.name = name;
  for (int c = 0; c < 3; c++) G.g[c] = g[c];
  e->grav.push_back(G);
  return ESP_OK;
  API_END
}

int esp_add_damping_group(esp_engine* e, int type, const char* name, double viscosity,
                          int max_iterations)
{
  API_BEGIN
  require_setup(e, "createInteractionGroup");
  static const char* kTypeName[4] = {"Damping", "RotDamping", "LocalDamping", "RotLocalDamping"};
  if (type < 0 || type > 3) throw EspError(ESP_ERR_ARG, "unknown damping type");
  if ((type == ESP_ROT_DAMPING || type == ESP_ROT_LOCAL_DAMPING) && !e->rot)
    throw EspError(ESP_ERR_ARG,
                   std::string("Trying to setup Damping of unknown type: ") + kTypeName[type]);
  if (!name) throw EspError(ESP_ERR_ARG, "no damping group name");
  const std::string key =
      (type == ESP_DAMPING || type == ESP_ROT_DAMPING) ? std::string(kTypeName[type]) : std::string(name);
  for (auto& D : e->damp)
    if (D.key == key) return ESP_OK;  // the map keeps its first entry
  if (e->damp.size() >= ESP_MAX_DAMPING) throw EspError(ESP_ERR_CAPACITY, "more than 4 damping groups");
  Damping D;
  D.type = type;
  D.key = key;
  D.visc = viscosity;
  D.maxit = max_iterations;
  e->damp.push_back(D);
  std::sort(e->damp.begin(), e->damp.end(),
            [](const Damping& a, const Damping& b) { return a.key < b.key; });
  return ESP_OK;
  API_END
}

int esp_move_wall_by(esp_engine* e, const char* name, const double d[3])
{
  API_BEGIN
  const int w = find_wall(e, name ? name : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall '") + (name ? name : "") + "'");
  for (int c = 0; c < 3; c++) e->walls[w].o[c] = e->walls[w].o[c] + d[c];
  return ESP_OK;
  API_END
}

namespace {
// applyForceToWall (LsmMpiPy.cpp:1208 -> TSubLattice::applyForceToWall,
// SubLattice.hpp:2935-2950) for wall interaction group k (name order):
// CEWallInteractionGroup::applyForce (EWallInteractionGroup.hpp:93-164) or
// CBWallInteractionGroup::applyForce (BWallInteractionGroup.hpp:71-145).  The
// candidates come from the device; the iteration runs here over them in the
// reference's list order (particle creation order), with its sums from zero.
// a wall candidate: creation rank, flags, position, initial position
void servo_iterate(esp_engine* e, int k, const double Fa[3], const WallCand* cs, int m);

struct CandRange {
  const WallCand* p;
  int m;
  const WallCand* begin() const { return p; }
  const WallCand* end() const { return p + m; }
};

void apply_force_to_wall(esp_engine* e, int k, const double Fa[3])
{
  const WallGroup& G = e->wg[e->word[k]];
  int m = 0;
  if (e->started && e->n > 0) {
    // the candidates (the group's wall-membership bit) in particle-list
    // (creation) order: marked by order on the device, packed at the scanned
    // offsets, so the host receives them sorted, in one copy
    const int no = (int)e->next_order;
    if (!e->ord_valid) {
      std::vector<int> os((size_t)std::max(no, 1), -1);
      for (int s = 0; s < e->n; s++) os[(size_t)e->h_order[s]] = s;
      e->ord_slot.reserve((size_t)std::max(no, 1));
      h2d(e, e->ord_slot.p, os.data(), (size_t)no);
      sync(e);  // the copy reads os, which ends with this block
      e->ord_valid = true;
    }
    e->wc_mark.reserve((size_t)no + 1);
    e->wc_off.reserve((size_t)no + 1);
    launch_wall_mark_ordered(e->stream, no, e->ord_slot.p, e->flags.p, k, e->wc_mark.p);
    HIPCHK(hipGetLastError());
    m = scan_total(e, e->wc_mark.p, e->wc_off.p, no);
    const size_t c = (size_t)std::max(m, 1);
    e->wc_cand.reserve(c);
    e->hw_cand.reserve(c);
    launch_wall_pack_ordered(e->stream, no, e->ord_slot.p, e->wc_mark.p, e->wc_off.p, e->view(),
                             G.bonded ? e->init.p : nullptr, e->wc_cand.p);
    HIPCHK(hipGetLastError());
    d2h(e, e->hw_cand.p, e->wc_cand.p, m);
    sync(e);
  }
  servo_iterate(e, k, Fa, e->hw_cand.p, m);
}

// the servo's iteration over the candidates in list order
void servo_iterate(esp_engine* e, int k, const double Fa[3], const WallCand* cs, int m)
{
  using Cand = WallCand;
  const WallGroup& G = e->wg[e->word[k]];
  Wall& W = e->walls[G.wall];
  auto inner = [&](const Cand& c) { return ((c.flags >> (PF_INNER_SHIFT + k)) & 1u) != 0; };
  auto tagged = [&](const Cand& c) {
    return (G.bonded || G.tagged) && ((c.flags >> (PF_BTAG_SHIFT + k)) & 1u) != 0;
  };
  const dv3 n = mk3(W.n[0], W.n[1], W.n[2]);
  // CElasticWallInteraction: k pi r (3-D, EWallInteraction.hpp:26-36); the
  // bonded form k pi r with its own constant (BWallInteraction.hpp:20-31)
  auto keff = [&](const Cand& c) { return (e->do2d ? 1.0 : 3.141592654 * c.pos.w) * G.k; };
  auto kbond = [&](const Cand& c) { return G.k * (e->do2d ? 1.0 : 3.1415926536 * c.pos.w); };
  // the iterations visit only the candidates their loops select, in list
  // order, with what does not change between iterations computed once (the
  // same operations as in the loops, so the sums are unchanged)
  struct Elastic {
    dv3 x;
    double r, ke;
    bool in;
  };
  struct Bonded {
    dv3 rel;  // position - initial position
    double kb;
  };
  auto dist = [&](const Elastic& c) { return dot3(sub3(c.x, mk3(W.o[0], W.o[1], W.o[2])), n); };
  auto elastic_force = [&](const Elastic& c) {  // CElasticWallInteraction::getForce
    dv3 f = mk3(0.0, 0.0, 0.0);
    const double d = dist(c);
    if (d < c.r) f = mul3(n, c.ke * (c.r - d));
    return mul3(f, -1.0);
  };
  auto elastic_of = [&](const Cand& c) {
    return Elastic{mk3(c.pos.x, c.pos.y, c.pos.z), c.pos.w, keff(c), inner(c)};
  };
  const dv3 F = mk3(Fa[0], Fa[1], Fa[2]);
  const dv3 O_f = div3(F, norm3(F));
  auto move = [&](double d) {
    const dv3 v = mul3(O_f, d);
    W.o[0] = W.o[0] + v.x;
    W.o[1] = W.o[1] + v.y;
    W.o[2] = W.o[2] + v.z;
  };
  std::vector<Elastic> el;
  el.reserve((size_t)m);
  if (!G.bonded) {
    // EWallInteractionGroup::applyForce: the stiffness sum over all (tagged)
    // candidates in contact, the force over the inner ones
    for (const Cand& c : CandRange{cs, m})
      if (!G.tagged || tagged(c)) el.push_back(elastic_of(c));
    int it = 0;
    double d, df, ef;
    do {
      double kk = 0.0;
      for (const Elastic& c : el)
        if (dist(c) < c.r) kk += c.ke;
      if (kk > 0) {
        dv3 Fl = mk3(0.0, 0.0, 0.0);
        for (const Elastic& c : el) {
          if (!c.in) continue;
          const dv3 f = elastic_force(c);
          Fl = add3(Fl, mul3(O_f, dot3(f, O_f)));
        }
        df = dot3(add3(F, Fl), O_f);
        d = df / kk;
        ef = df / norm3(F);
      } else {
        d = 1e-5;
        ef = 1;
      }
      move(d);
      it++;
    } while ((it < 50) && (ef > 1e-3));
    return;
  }
  // BWallInteractionGroup::applyForce: bonded candidates, then the unbonded
  // inner ones in contact, make the stiffness; the iterations sum the bonded
  // inner forces, then the unbonded inner ones
  std::vector<Bonded> bd;
  bd.reserve((size_t)m);
  double K = 0.0;
  for (const Cand& c : CandRange{cs, m})
    if (tagged(c) && inner(c)) {
      const double kb = kbond(c);
      K += kb;
      bd.push_back(Bonded{sub3(mk3(c.pos.x, c.pos.y, c.pos.z), mk3(c.ix, c.iy, c.iz)), kb});
    }
  for (const Cand& c : CandRange{cs, m})
    if (!tagged(c) && inner(c)) {
      el.push_back(elastic_of(c));
      if (dist(el.back()) < c.pos.w) K += el.back().ke;
    }
  if (!(K > 0)) throw EspError(ESP_ERR_NUMERIC, "wall group '" + G.name + "' has no stiffness to servo");
  int it = 0;
  double d;
  do {
    const dv3 wdi = mk3(W.o[0] - W.old[0], W.o[1] - W.old[1], W.o[2] - W.old[2]);
    dv3 Fl = mk3(0.0, 0.0, 0.0);
    for (const Bonded& c : bd) {
      const dv3 D = sub3(c.rel, wdi);
      const dv3 f = mul3(D, c.kb);
      Fl = add3(Fl, mul3(O_f, dot3(f, O_f)));
    }
    for (const Elastic& c : el) {
      const dv3 f = elastic_force(c);
      Fl = add3(Fl, mul3(O_f, dot3(f, O_f)));
    }
    d = dot3(add3(F, Fl), O_f) / K;
    move(d);
    it++;
  } while ((it < 10) && (std::fabs(d) > 10e-6));
}
}  // namespace

namespace {
// ---------------------------------------------------------------------------
// applyForceToWall without allocation: the candidates are staged in a
// file-scope device array (one copy per GPU, held until the process ends) and
// in file-scope host arrays, instead of buffers allocated per call.  Calls
// that overlap on one GPU (or in one process) race on these arrays.
// ---------------------------------------------------------------------------
constexpr size_t NA_CNT = 0;  // candidate counter (one int)
constexpr size_t NA_SLOT = 256;
constexpr size_t NA_FLG = NA_SLOT + sizeof(int) * ESP_NOALLOC_MAX_WALL_CANDIDATES;
constexpr size_t NA_POS = NA_FLG + sizeof(unsigned) * ESP_NOALLOC_MAX_WALL_CANDIDATES;
constexpr size_t NA_IPOS = NA_POS + sizeof(double4) * ESP_NOALLOC_MAX_WALL_CANDIDATES;
constexpr size_t NA_BYTES = NA_IPOS + sizeof(double4) * ESP_NOALLOC_MAX_WALL_CANDIDATES;
static_assert(NA_POS % 32 == 0 && NA_IPOS % 32 == 0, "double4 staging must stay aligned");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NA_MAX_DEVICES] = {};

// the device address of `scratch` on `dev` (the current device), looked up
// once per device; host code never passes the symbol itself to a kernel
void* scratch_on(int dev)
{
  if (dev < 0 || dev >= NA_MAX_DEVICES)
    throw EspError(ESP_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
      throw EspError(ESP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(rc));
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

int h_na_slot[ESP_NOALLOC_MAX_WALL_CANDIDATES];
unsigned h_na_flg[ESP_NOALLOC_MAX_WALL_CANDIDATES];
double4 h_na_pos[ESP_NOALLOC_MAX_WALL_CANDIDATES];
double4 h_na_ipos[ESP_NOALLOC_MAX_WALL_CANDIDATES];
WallCand h_na_cand[ESP_NOALLOC_MAX_WALL_CANDIDATES];
int h_na_idx[ESP_NOALLOC_MAX_WALL_CANDIDATES];

void apply_force_to_wall_noalloc(esp_engine* e, int k, const double Fa[3])
{
  const WallGroup& G = e->wg[e->word[k]];
  int m = 0;
  if (e->started && e->n > 0) {
    unsigned char* base = static_cast<unsigned char*>(scratch_on(e->cfg.device));
    int* cnt = reinterpret_cast<int*>(base + NA_CNT);
    int* slot = reinterpret_cast<int*>(base + NA_SLOT);
    unsigned* flg = reinterpret_cast<unsigned*>(base + NA_FLG);
    double4* pos = reinterpret_cast<double4*>(base + NA_POS);
    double4* ipos = reinterpret_cast<double4*>(base + NA_IPOS);
    launch_fill_int(e->stream, 1, cnt, 0);  // the last call left its count
    launch_wall_cands(e->stream, e->view(), k, G.bonded ? e->init.p : nullptr,
                      ESP_NOALLOC_MAX_WALL_CANDIDATES, cnt, slot, flg, pos, ipos);
    HIPCHK(hipGetLastError());
    d2h(e, &m, cnt, 1);
    sync(e);
    if (m > ESP_NOALLOC_MAX_WALL_CANDIDATES)
      throw EspError(ESP_ERR_CAPACITY,
                     "wall group '" + G.name + "' has " + std::to_string(m) +
                         " candidate particles; the allocation-free servo holds at most " +
                         std::to_string(ESP_NOALLOC_MAX_WALL_CANDIDATES));
    d2h(e, h_na_slot, slot, m);
    d2h(e, h_na_flg, flg, m);
    d2h(e, h_na_pos, pos, m);
    if (G.bonded) d2h(e, h_na_ipos, ipos, m);
    sync(e);
    // the candidates arrive in any order: sorted into particle-list order
    for (int c = 0; c < m; c++) h_na_idx[c] = c;
    std::sort(h_na_idx, h_na_idx + m, [&](int a, int b) {
      return e->h_order[h_na_slot[a]] < e->h_order[h_na_slot[b]];
    });
    for (int i = 0; i < m; i++) {
      const int c = h_na_idx[i];
      const double4 x0 = G.bonded ? h_na_ipos[c] : make_double4(0, 0, 0, 0);
      h_na_cand[i] = WallCand{h_na_pos[c], x0.x, x0.y, x0.z, h_na_flg[c], 0u};
    }
  }
  servo_iterate(e, k, Fa, h_na_cand, m);
}
}  // namespace

int esp_apply_force_to_wall_noalloc(esp_engine* e, const char* name, const double F[3])
{
  API_BEGIN
  if (!name || !F) throw EspError(ESP_ERR_ARG, "bad arguments");
  if (e->word.size() != e->wg.size()) refresh_orders(e);  // fixed once the run started
  int k = -1;
  for (size_t i = 0; i < e->word.size(); i++)
    if (e->wg[e->word[i]].name == name) k = (int)i;
  if (k < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall interaction group '") + name + "'");
  HIPCHK(hipSetDevice(e->cfg.device));
  apply_force_to_wall_noalloc(e, k, F);
  return ESP_OK;
  API_END
}

int esp_apply_force_to_wall(esp_engine* e, const char* name, const double F[3])
{
  API_BEGIN
  if (!name || !F) throw EspError(ESP_ERR_ARG, "bad arguments");
  refresh_orders(e);
  int k = -1;
  for (size_t i = 0; i < e->word.size(); i++)
    if (e->wg[e->word[i]].name == name) k = (int)i;
  if (k < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall interaction group '") + name + "'");
  apply_force_to_wall(e, k, F);
  return ESP_OK;
  API_END
}

int esp_set_wall_step_motion(esp_engine* e, const char* name, const double d[3])
{
  API_BEGIN
  const int w = find_wall(e, name ? name : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, std::string("unknown wall '") + (name ? name : "") + "'");
  for (int c = 0; c < 3; c++) e->walls[w].step[c] = d[c];
  e->walls[w].has_motion = (d[0] != 0.0 || d[1] != 0.0 || d[2] != 0.0);
  return ESP_OK;
  API_END
}

int esp_run_time_steps(esp_engine* e, int nsteps)
{
  API_BEGIN
  if (nsteps < 0) throw EspError(ESP_ERR_ARG, "negative step count");
  if (nsteps == 0) return ESP_OK;
  HIPCHK(hipSetDevice(e->cfg.device));
  run_steps(e, nsteps);
  return ESP_OK;
  API_END
}

int esp_get_num_particles(esp_engine* e) { return e->n; }

int esp_get_time_step(esp_engine* e, long long* out)
{
  API_BEGIN
  if (!out) throw EspError(ESP_ERR_ARG, "bad arguments");
  *out = e->t;
  return ESP_OK;
  API_END
}

int esp_get_stats(esp_engine* e, esp_stats* o)
{
  API_BEGIN
  if (e->started) fold_events(e);
  memset(o, 0, sizeof *o);
  o->time_step = e->t;
  o->num_rebuilds = e->timestamp;
  o->num_particles = e->n;
  for (auto& G : e->bg) o->num_bonds += bond_live_count(e, G);
  for (auto& G : e->dg) o->num_dynamic += G.n + dyn_appended(e, G);
  o->num_candidates = e->ncand;
  o->num_broken = (long long)e->breaks.size() / 3;
  o->num_host_phases = e->num_host_phases;
  return ESP_OK;
  API_END
}

int esp_get_particle_ids(esp_engine* e, int* out)
{
  API_BEGIN
  std::vector<int> order = id_order(e);
  for (int k = 0; k < e->n; k++) out[k] = e->h_id[order[k]];
  return ESP_OK;
  API_END
}

int esp_get_particle_data(esp_engine* e, int field, double* out)
{
  API_BEGIN
  const int n = e->n;
  std::vector<int> order = id_order(e);
  if (!e->started) {
    for (int k = 0; k < n; k++) {
      const int i = order[k];
      const double r = e->h_rad[i];
      switch (field) {
        case ESP_FIELD_POS:
        case ESP_FIELD_OLDPOS:
          for (int c = 0; c < 3; c++) out[3 * k + c] = e->h_pos[3 * i + c];
          break;
        case ESP_FIELD_VEL:
          for (int c = 0; c < 3; c++) out[3 * k + c] = e->h_vel[3 * i + c];
          break;
        case ESP_FIELD_ANGVEL:
        case ESP_FIELD_FORCE:
        case ESP_FIELD_MOMENT:
          for (int c = 0; c < 3; c++) out[3 * k + c] = 0.0;
          break;
        case ESP_FIELD_QUAT:
          out[4 * k] = 1.0;
          out[4 * k + 1] = out[4 * k + 2] = out[4 * k + 3] = 0.0;
          break;
        case ESP_FIELD_RAD: out[k] = r; break;
        case ESP_FIELD_MASS: out[k] = e->h_mass[i]; break;
        default: throw EspError(ESP_ERR_ARG, "unknown field");
      }
    }
    return ESP_OK;
  }
  HIPCHK(hipSetDevice(e->cfg.device));
  std::vector<double4> buf(std::max(n, 1));
  const double4* src = nullptr;
  int sel = 0;  // 0 xyz, 1 quat, 2 w
  switch (field) {
    case ESP_FIELD_POS: src = e->pos.p; break;
    case ESP_FIELD_VEL: src = e->vel.p; break;
    case ESP_FIELD_ANGVEL: src = e->rot ? e->ang.p : nullptr; break;
    case ESP_FIELD_QUAT: src = e->rot ? e->quat.p : nullptr; sel = 1; break;
    case ESP_FIELD_FORCE: src = e->frc.p; break;
    case ESP_FIELD_MOMENT: src = e->rot ? e->mom.p : nullptr; break;
    case ESP_FIELD_RAD: src = e->pos.p; sel = 2; break;
    case ESP_FIELD_MASS: src = e->old.p; sel = 2; break;
    case ESP_FIELD_OLDPOS: src = e->old.p; break;
    default: throw EspError(ESP_ERR_ARG, "unknown field");
  }
  if (src) {
    d2h(e, buf.data(), src, n);
    sync(e);
  }
  for (int k = 0; k < n; k++) {
    const int i = order[k];
    if (!src) {
      if (sel == 1) { out[4 * k] = 1.0; out[4 * k + 1] = out[4 * k + 2] = out[4 * k + 3] = 0.0; }
      else for (int c = 0; c < 3; c++) out[3 * k + c] = 0.0;
      continue;
    }
    const double4 v = buf[i];
    if (sel == 0) { out[3 * k] = v.x; out[3 * k + 1] = v.y; out[3 * k + 2] = v.z; }
    else if (sel == 1) { out[4 * k] = v.x; out[4 * k + 1] = v.y; out[4 * k + 2] = v.z; out[4 * k + 3] = v.w; }
    else out[k] = v.w;
  }
  return ESP_OK;
  API_END
}

int esp_get_wall_position(esp_engine* e, const char* name, double out[3])
{
  API_BEGIN
  const int w = find_wall(e, name ? name : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, "unknown wall");
  for (int c = 0; c < 3; c++) out[c] = e->walls[w].o[c];
  return ESP_OK;
  API_END
}

int esp_get_wall_force(esp_engine* e, const char* name, double out[3])
{
  API_BEGIN
  const int w = find_wall(e, name ? name : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, "unknown wall");
  for (int c = 0; c < 3; c++) out[c] = e->walls[w].force[c];
  return ESP_OK;
  API_END
}

int esp_get_wall_force_history(esp_engine* e, const char* name, int n, double* out)
{
  API_BEGIN
  const int w = find_wall(e, name ? name : "");
  if (w < 0) throw EspError(ESP_ERR_ARG, "unknown wall");
  const int have = (int)e->walls[w].hist.size() / 3;
  const int m = std::min(n, have);
  if (out && m > 0)
    memcpy(out, e->walls[w].hist.data() + 3 * (size_t)(have - m), sizeof(double) * 3 * m);
  return m;
  API_END
}

int esp_get_interaction_count(esp_engine* e, const char* name)
{
  API_BEGIN
  const int d = find_dg(e, name ? name : "");
  if (d >= 0) return e->dg[d].n + dyn_appended(e, e->dg[d]);
  const int b = find_bg(e, name ? name : "");
  if (b >= 0) return bond_live_count(e, e->bg[b]);
  return -1;
  API_END
}

int esp_get_interactions(esp_engine* e, const char* name, int* id1, int* id2)
{
  API_BEGIN
  const int d = find_dg(e, name ? name : "");
  std::vector<long long> k;
  if (d >= 0) {
    dyn_view(e, e->dg[d], k, nullptr);
  } else {
    const int b = find_bg(e, name ? name : "");
    if (b < 0) throw EspError(ESP_ERR_ARG, "unknown interaction group");
    BondGroup& G = e->bg[b];
    const std::vector<char> alive = bond_alive(e, G);
    for (int i = 0; i < G.n; i++)
      if (alive[i]) k.push_back(pkey(G.hid1[i], G.hid2[i]));
    std::sort(k.begin(), k.end());
  }
  for (size_t i = 0; i < k.size(); i++) {
    id1[i] = (int)(k[i] >> 32);
    id2[i] = (int)(k[i] & 0xffffffffLL);
  }
  return (int)k.size();
  API_END
}

int esp_set_interaction_group_tags(esp_engine* e, const char* name, int tag1, int mask1,
                                   int tag2, int mask2)
{
  API_BEGIN
  require_setup(e, "createInteractionGroupTagged");
  const int d = find_dg(e, name ? name : "");
  if (d < 0) throw EspError(ESP_ERR_ARG, "not a dynamic interaction group");
  DynGroup& G = e->dg[d];
  G.tagged = 1;
  if (tag1 <= tag2) {
    G.tag1 = tag1; G.mask1 = mask1; G.tag2 = tag2; G.mask2 = mask2;
  } else {
    G.tag1 = tag2; G.mask1 = mask2; G.tag2 = tag1; G.mask2 = mask1;
  }
  return ESP_OK;
  API_END
}

int esp_track_bond_criteria(esp_engine* e)
{
  API_BEGIN
  require_setup(e, "createFieldSaver(breaking_criterion)");
  e->track_crit = true;
  return ESP_OK;
  API_END
}

int esp_get_bond_criteria(esp_engine* e, const char* name, int* id1, int* id2, double* crit)
{
  API_BEGIN
  const int b = find_bg(e, name ? name : "");
  if (b < 0) throw EspError(ESP_ERR_ARG, "not a bonded interaction group");
  if (!e->track_crit) throw EspError(ESP_ERR_STATE, "bond criteria are not tracked");
  BondGroup& G = e->bg[b];
  const std::vector<char> alive = bond_alive(e, G);
  std::vector<double> c(std::max(G.n, 1), 0.0);
  std::vector<int> rank(std::max(G.n, 1));
  if (e->started && G.n > 0) {
    d2h(e, c.data(), e->crit.p + G.bond_off, G.n);
    d2h(e, rank.data(), G.rank.p, G.n);
    sync(e);
  } else {
    for (int i = 0; i < G.n; i++) rank[i] = i;
  }
  std::vector<int> order;
  for (int i = 0; i < G.n; i++)
    if (alive[i]) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return rank[x] < rank[y]; });
  if (id1 && id2 && crit)
    for (size_t k = 0; k < order.size(); k++) {
      id1[k] = G.hid1[order[k]];
      id2[k] = G.hid2[order[k]];
      crit[k] = c[order[k]];
    }
  return (int)order.size();
  API_END
}

int esp_get_friction_state(esp_engine* e, const char* name, double* ffric, int* flg)
{
  API_BEGIN
  const int d = find_dg(e, name ? name : "");
  if (d < 0 || e->dg[d].kind != 3) throw EspError(ESP_ERR_ARG, "not a RotFriction group");
  std::vector<long long> k;
  std::vector<double4> f;
  dyn_view(e, e->dg[d], k, &f);
  for (size_t i = 0; i < f.size(); i++) {
    ffric[3 * i] = f[i].x;
    ffric[3 * i + 1] = f[i].y;
    ffric[3 * i + 2] = f[i].z;
    flg[i] = (int)f[i].w & 3;  // bit 0 slipping, bit 1 touching
  }
  return (int)f.size();
  API_END
}

int esp_save_checkpoint(esp_engine* e, const char* prefix, int num_time_steps, int precision)
{
  API_BEGIN
  if (!prefix || !*prefix) throw EspError(ESP_ERR_ARG, "empty checkpoint file prefix");
  if (precision < 1 || precision > 17) throw EspError(ESP_ERR_ARG, "precision must be 1..17");
  if (e->started) {
    HIPCHK(hipSetDevice(e->cfg.device));
    // the reference searches neighbours before every checkpoint
    // (LatticeMaster.cpp:2697-2700); the next step then updates every group.
    // The search may erase particles, whose remap of the dynamic groups needs
    // key-sorted lists without appendices.  Dead bonds stay (the erasure drops
    // them), so bond indices change only where the erasure re-indexes the last
    // step's pending breaks with them.
    fold_events(e);
    fold_appendices(e);
    rebuild(e, e->t);
    e->forced_search = true;
    e->pending_host = true;
    e->pending_reason = 0;
  }
  cp_write(e, prefix, num_time_steps, precision);
  return ESP_OK;
  API_END
}

int esp_load_checkpoint(esp_engine* e, const char* info_file)
{
  API_BEGIN
  require_setup(e, "loadCheckPoint");
  if (!info_file) throw EspError(ESP_ERR_ARG, "no checkpoint file");
  if (e->n > 0) throw EspError(ESP_ERR_STATE, "loadCheckPoint needs a model without particles");
  cp_load(e, info_file);
  return ESP_OK;
  API_END
}

int esp_get_particle_stress(esp_engine* e, double* out)
{
  API_BEGIN
  if (!e->started || e->ss != 9)
    throw EspError(ESP_ERR_STATE, "particle stresses are tracked only with a BrittleBeamSC group, "
                                  "after the first time step");
  HIPCHK(hipSetDevice(e->cfg.device));
  std::vector<double> s((size_t)std::max(e->n, 1) * 9);
  d2h(e, s.data(), e->sig.p, (size_t)e->n * 9);
  sync(e);
  std::vector<int> order = id_order(e);
  for (int k = 0; k < e->n; k++)
    for (int c = 0; c < 9; c++) out[9 * k + c] = s[9 * (size_t)order[k] + c];
  return ESP_OK;
  API_END
}

int esp_get_neighbour_pairs(esp_engine* e, int* id1, int* id2)
{
  API_BEGIN
  if (id1 && id2 && e->ncand > 0) {
    std::vector<long long> k(e->ncand);
    d2h(e, k.data(), e->cand_key.p, e->ncand);
    sync(e);
    for (int i = 0; i < e->ncand; i++) k[i] = id_key(e, k[i]);
    std::sort(k.begin(), k.end());
    for (int i = 0; i < e->ncand; i++) {
      id1[i] = (int)(k[i] >> 32);
      id2[i] = (int)(k[i] & 0xffffffffLL);
    }
  }
  return e->ncand;
  API_END
}

int esp_track_pair_events(esp_engine* e, const char* name)
{
  API_BEGIN
  require_setup(e, "trackPairEvents");
  const int d = find_dg(e, name ? name : "");
  if (d < 0) throw EspError(ESP_ERR_ARG, std::string("no dynamic interaction group '") + (name ? name : "") + "'");
  e->dg[d].track = 1;
  return ESP_OK;
  API_END
}

int esp_get_pair_events(esp_engine* e, const char* name, int* step, int* id1, int* id2, int* kind)
{
  API_BEGIN
  const int d = find_dg(e, name ? name : "");
  if (d < 0) throw EspError(ESP_ERR_ARG, std::string("no dynamic interaction group '") + (name ? name : "") + "'");
  DynGroup& G = e->dg[d];
  if (!G.track) throw EspError(ESP_ERR_STATE, std::string("pair events of '") + name + "' are not tracked");
  harvest_appendix(e, G);
  const int n = (int)G.ev.size();
  if (step && id1 && id2 && kind) {
    std::vector<std::array<int, 4>> v(n);
    for (int i = 0; i < n; i++) v[i] = {G.ev[i].x, G.ev[i].y, G.ev[i].z, G.ev[i].w};
    std::sort(v.begin(), v.end());
    for (int i = 0; i < n; i++) { step[i] = v[i][0]; id1[i] = v[i][1]; id2[i] = v[i][2]; kind[i] = v[i][3]; }
  }
  return n;
  API_END
}

int esp_get_bond_breaks(esp_engine* e, int* step, int* id1, int* id2)
{
  API_BEGIN
  if (e->started) fold_events(e);
  const int n = (int)e->breaks.size() / 3;
  if (step && id1 && id2) {
    std::vector<std::array<int, 3>> v(n);
    for (int i = 0; i < n; i++) v[i] = {e->breaks[3 * i], e->breaks[3 * i + 1], e->breaks[3 * i + 2]};
    std::sort(v.begin(), v.end());
    for (int i = 0; i < n; i++) { step[i] = v[i][0]; id1[i] = v[i][1]; id2[i] = v[i][2]; }
  }
  return n;
  API_END
}

}  // extern "C"

namespace {
// Consistency of the device bookkeeping that queued steps index with: every
// per-particle appendix count, flagged-list entry, event count, pair, key and
// force-record slot inside its buffer (and a record's CSR slots inside its
// particles' runs).  When the next step is a host bookkeeping step (a halted
// queue, a checkpoint's search), that step lays the records and the CSR out
// anew from the lists: then only the lists are checked.  Returns an empty
// string when consistent.
std::string check_state(esp_engine* e)
{
  if (!e->started) return "";
  HIPCHK(hipSetDevice(e->cfg.device));
  sync(e);
  const int n = e->n;
  const bool csr = !e->pending_host;  // records and CSR the next step indexes with
  std::ostringstream err;
  auto fail = [&](const std::string& m) {
    if (err.tellp() < 2000) err << m << "; ";
  };
  std::vector<int> xcnt(std::max(n, 1)), off(n + 1);
  UpdState st;
  d2h(e, xcnt.data(), e->xcnt.p, n);
  d2h(e, off.data(), e->inc_off.p, n + 1);
  d2h(e, &st, e->upd.p, 1);
  std::vector<int2> so(std::max(e->nrec, 1));
  d2h(e, so.data(), e->slot_of.p, e->nrec);
  sync(e);
  for (int i = 0; i < n; i++)
    if (xcnt[i] < 0 || xcnt[i] > ESP_XSLOTS) {
      fail("appendix slot count " + std::to_string(xcnt[i]) + " at slot " + std::to_string(i));
      break;
    }
  for (int i = 0; i < n && csr; i++)
    if (off[i] > off[i + 1]) {
      fail("CSR offsets decrease at slot " + std::to_string(i));
      break;
    }
  const long long nslot = (long long)e->xbase + (long long)n * ESP_XSLOTS;
  if (csr && off[n] != e->xbase) fail("CSR total differs from the appendix base");
  if (st.fcount < 0 || st.fcount > n) fail("flagged count " + std::to_string(st.fcount));
  if (st.evt_count < 0 || st.evt_count > e->broken_cap)
    fail("event count " + std::to_string(st.evt_count));
  if (st.fcount > 0 && st.fcount <= n) {
    std::vector<int> fl(st.fcount);
    d2h(e, fl.data(), e->flist.p, st.fcount);
    sync(e);
    for (int v : fl)
      if (v < 0 || v >= n) {
        fail("flagged particle slot " + std::to_string(v));
        break;
      }
  }
  auto in_run = [&](int slot, int p) { return slot >= off[p] && slot < off[p + 1]; };
  for (int gi : e->bord) {
    BondGroup& G = e->bg[gi];
    std::vector<int2> ij(std::max(G.n, 1));
    d2h(e, ij.data(), G.ij.p, G.n);
    sync(e);
    for (int b = 0; b < G.n; b++) {
      const int2 q = ij[b];
      if (q.x == -1 && q.y == -1) continue;  // dead until the next compaction
      const int2 sl = csr ? so[G.rec_off + b] : make_int2(0, 0);
      if (q.x < 0 || q.x >= n || q.y < 0 || q.y >= n || q.x == q.y ||
          (csr && (!in_run(sl.x, q.x) || !in_run(sl.y, q.y)))) {
        fail("bond " + std::to_string(b) + " of '" + G.name + "'");
        break;
      }
    }
  }
  for (int gi : e->dord) {
    DynGroup& G = e->dg[gi];
    int na[2] = {0, 0};
    d2h(e, na, G.napp.p, 2);
    sync(e);
    if (na[0] < 0 || na[0] > ESP_APP_CAP || na[1] < 0 || na[1] > na[0]) {
      fail("appendix counts of '" + G.name + "'");
      continue;
    }
    const int tot = G.n + na[0];
    std::vector<int2> ij(std::max(tot, 1));
    std::vector<long long> key(std::max(tot, 1));
    d2h(e, ij.data(), G.ij.p, tot);
    d2h(e, key.data(), G.key.p, tot);
    sync(e);
    for (int c = 0; c < tot; c++) {
      const int2 q = ij[c];
      const bool app = c >= G.n;
      const int2 sl = !csr ? make_int2(0, 0)
                           : app ? so[G.app_rec_off + (c - G.n)] : so[G.rec_off + c];
      bool ok = q.x >= 0 && q.x < n && q.y >= 0 && q.y < n && q.x != q.y && key[c] == pkey(q.x, q.y);
      if (ok && !app && csr) ok = in_run(sl.x, q.x) && in_run(sl.y, q.y);
      if (ok && app && csr)
        ok = sl.x >= e->xbase + ESP_XSLOTS * q.x && sl.x < e->xbase + ESP_XSLOTS * (q.x + 1) &&
             sl.y >= e->xbase + ESP_XSLOTS * q.y && sl.y < e->xbase + ESP_XSLOTS * (q.y + 1) &&
             sl.y < nslot;
      if (!ok) {
        fail("pair " + std::to_string(c) + " of '" + G.name + "'" + (app ? " (appendix)" : ""));
        break;
      }
    }
  }
  return err.str();
}
}  // namespace

extern "C" {

int esp_check_state(esp_engine* e)
{
  API_BEGIN
  const std::string m = check_state(e);
  if (!m.empty()) throw EspError(ESP_ERR_STATE, "device bookkeeping inconsistent: " + m);
  return ESP_OK;
  API_END
}

int esp_get_rebuild_steps(esp_engine* e, int* steps)
{
  API_BEGIN
  if (steps && !e->rebuild_steps.empty())
    memcpy(steps, e->rebuild_steps.data(), sizeof(int) * e->rebuild_steps.size());
  return (int)e->rebuild_steps.size();
  API_END
}

int esp_set_kernel_timing(esp_engine* e, int enable)
{
  API_BEGIN
  e->timing = enable != 0;
  e->tl.clear();
  e->ev_used = 0;
  for (int k = 0; k < ESP_KCLASS_COUNT; k++) { e->k_ms[k] = 0; e->k_n[k] = 0; }
  return ESP_OK;
  API_END
}

int esp_get_kernel_times(esp_engine* e, double* ms, long long* launches)
{
  API_BEGIN
  for (int k = 0; k < ESP_KCLASS_COUNT; k++) {
    if (ms) ms[k] = e->k_ms[k];
    if (launches) launches[k] = e->k_n[k];
    e->k_ms[k] = 0;
    e->k_n[k] = 0;
  }
  return ESP_OK;
  API_END
}

}  // extern "C"
