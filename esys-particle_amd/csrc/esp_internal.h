// esp_internal.h - device data layout and kernel launchers of the engine.
//
// Layout in HBM (structure of arrays; every per-particle field is one array of
// double4 so that one 32-byte load fetches a field for a gathered neighbour):
//   pos[i]    = (x, y, z, radius)
//   vel[i]    = (vx, vy, vz, 1/mass)
//   ang[i]    = (wx, wy, wz, 1/inertia)        rotational engines only
//   quat[i]   = (s, x, y, z)                   rotational engines only
//   old[i]    = (x, y, z, mass)                position at the last neighbour search
//   frc[i], mom[i]                             total force / moment of the last step
//   flags[i]  = dynamic | rotational<<1 | bond-broken flag<<2 |
//               wall-group membership<<8 (8 bits) | wall-group inner<<16 (8 bits) |
//               tag matches bonded wall group<<24 (8 bits)
// Interactions are structure-of-arrays per group (pair slots int2, history
// double4).  Every interaction has a record index; at each rebuild a
// per-particle CSR orders a particle's interactions for summation and slot_of
// maps each record to its two CSR positions.  Per step an interaction writes
// each particle's share (force, moment: 6 doubles; with particle stresses
// tracked also the contact point relative to the particle: 9 doubles) into that
// particle's slot, and each particle sums its contiguous run of slots in order -
// the deterministic, atomic-free force reduction.
// Shear histories are double-buffered: a step reads one buffer and writes the
// other, so a queued step that has to be redone after host bookkeeping still
// finds its input.  Flag bit 2 of a history marks "zero in both buffers".
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/esp.h"

#define ESP_MAX_WALL_GROUPS 8
#define ESP_MAX_GRAVITY 4
#define ESP_MAX_DAMPING 4
#define ESP_MAX_EXCL 4

#define PF_DYN 1u
#define PF_ROT 2u
#define PF_FLAG 4u
#define PF_WALL_SHIFT 8
#define PF_INNER_SHIFT 16
#define PF_BTAG_SHIFT 24  // bit 24 + w: tag matches bonded wall group w (static)

// Status of one time step, written by that step's kernels.
struct StepStatus {
  // bits of max |x - x_old|^2 when it exceeds (verletDist/2)^2, else 0
  // (non-negative doubles order as integers)
  unsigned long long maxdisp2;
  int nbroken;                  // bonds whose break criterion held in this step
  int pad;
};

// Per-step guard shared by the kernels of one queued step.
struct StepCtl {
  const StepStatus* prev;   // status of the previous step
  StepStatus* cur;          // status of this step
  int* halted;              // 0, or 1 + index of the first step that needs the host
  unsigned long long thresh;  // bits of (verletDist/2)^2
  int step_id;              // 1 + index of this step in its batch
  int guarded;              // 0 for steps launched after host bookkeeping
  int* reason;              // why the queue halted (UpdState::halt_reason)
};

struct GridParams {
  double x0, y0, z0, dim;
  int gx, gy, gz;  // global index of the table's minimum corner (-1)
  int nx, ny, nz;
  // circular x: no padding cells in x (gx = 0), the stencil wraps and the
  // period is L = nx * dim (0: not circular)
  int circ_x;
  double period;
};

struct PView {
  double4* pos;
  double4* vel;
  double4* ang;
  double4* quat;
  double4* old;
  double4* frc;
  double4* mom;
  unsigned* flags;
  int* id;
  int n;
};

struct WallParam {
  double o[3];
  double nrm[3];
  double k;
  int bonded;      // CBWallInteractionGroup: tagged candidates are bonded to the wall
  int tagged;      // CTaggedEWallInteractionGroup: only tagged candidates interact
  double disp[3];  // CWall::getTotalDisplacement: origin - origin at creation
};

struct GatherParams {
  const int* inc_off;  // CSR runs of slots
  const int* xcnt;     // appendix slots in use per particle (ESP_XSLOTS each, from xbase)
  int xbase;
  GridParams grid;     // wall candidates are refreshed after bond breaks
  const double* slots;
  const unsigned char* sact;  // per slot: nonzero contents
  double dt;
  int rot, do2d;
  int nwg;
  WallParam wall[ESP_MAX_WALL_GROUPS];
  int ngrav;
  double g[ESP_MAX_GRAVITY][3];
  double* wall_part;  // this step's slot: [block][nwg][3]
  double* sig;        // stress sums [particle][9] (slots of 9 doubles only)
  int* n_app;         // appendix counters of the incrementally updated group, or nullptr
  int store_fm;       // store the total force and moment (the last step before the host reads them)
  // damping groups, applied last in the reference's map order of their keys
  int ndamp;
  int damp_type[ESP_MAX_DAMPING];  // ESP_DAMPING, ESP_ROT_DAMPING, ...
  double damp_visc[ESP_MAX_DAMPING];
  int damp_maxit[ESP_MAX_DAMPING];
  const double* inert;  // moment of inertia per particle (rotational damping)
  const double4* init;  // position at creation per particle (bonded walls)
};

// Open-addressing hash sets of pair keys (a << 32 | b): live bonds of a bonded
// group keyed by particle ids (exclusion lookups), stored pairs of a dynamic
// group keyed by particle slots (isIn).  A pair is oriented (lower id, higher
// id); dynamic lists are sorted by the slot key of that orientation.
#define HS_EMPTY 0xFFFFFFFFFFFFFFFFull
#define HS_TOMB 0xFFFFFFFFFFFFFFFEull
struct HashRef {
  unsigned long long* t;
  unsigned mask;
};

// what a dynamic group refuses besides stored pairs: pairs of an excluding
// bonded group (isExcluded) and, for groups created with
// createInteractionGroupTagged, pairs whose tags do not match
// (pi_storage_ed_t.hpp:55-70; tag1 <= tag2)
struct ExclSet {
  HashRef h[ESP_MAX_EXCL];
  int count;
  int tagged, tag1, mask1, tag2, mask2;
  const int* tag;  // per particle slot
};

#define ESP_MAX_BOND_GROUPS 8
// bonded groups as the break-application kernel sees them
struct BondSet {
  int count;
  int ss;               // doubles per slot
  double* slots;
  unsigned char* sact;  // per slot: nonzero contents
  const int2* slot_of;  // record -> its two slots
  int off[ESP_MAX_BOND_GROUPS];  // global bond index of the group's first bond
  int n[ESP_MAX_BOND_GROUPS];
  int rec_off[ESP_MAX_BOND_GROUPS];
  int2* ij[ESP_MAX_BOND_GROUPS];
  HashRef hs[ESP_MAX_BOND_GROUPS];
};

// Device-side bookkeeping state shared by the update kernels.
struct UpdState {
  int fcount;       // particles in the flagged list
  int evt_count;    // bond-break events logged on the device
  int halt_reason;  // 1 rebuild due, 2 incremental update does not fit
  int pad;
};

// Appendix of a dynamic group: pairs added between neighbour-table rebuilds
// are stored after the key-sorted main list, with their force records in a
// separate region and per-particle overflow slots in the CSR.
#define ESP_APP_CAP 16384
#define ESP_NEW_CAP 4096
#define ESP_XSLOTS 4
struct DynAppend {
  int2* slot_of;     // record -> slots; appended records get appendix slots
  int xbase;         // first appendix slot
  int2* ij;          // whole list: main [0, n_main), appendix after it
  long long* key;
  double4* ff;       // shear history read by this step (history groups) or nullptr
  int* n_app;        // device counters: [0] pairs appended, [1] pairs the force blocks cover
  int n_main;
  int rec_off;       // records of the main list
  int app_rec_off;   // records of the appendix
  int is_hist;       // carries history / checks presence
  HashRef present;   // stored pairs (history groups)
  ExclSet ex;
};

// ---- per-step kernels ----
// One block segment per interaction group in the fused force-law launch.
#define ESP_SEG_NROT_ELASTIC 1
#define ESP_SEG_ROT_ELASTIC 2
#define ESP_SEG_ROT_FRICTION 3
#define ESP_SEG_ROT_BOND 4
#define ESP_MAX_FORCE_SEGS 16
struct ForceSeg {
  int kind;
  int blk0;         // first block of the segment
  int n;            // bonds, or pairs of a dynamic group's main list
  int rec_off;      // records of the bonds / of the main list
  int app_rec_off;  // records of the appendix (dynamic groups)
  int bond_off;     // global index of the first bond (break list entries)
  int scaling, meanR;
  int sc;            // bonded group broken by the stress criterion (no force test)
  int rigid;         // RotFriction rigid=True (calcRigidBodyForces)
  const int* n_app;  // appended pairs the force blocks cover, device count (dynamic groups)
  const int2* ij;
  double4* ff;       // shear history read by this step (RotFriction)
  double4* ff_out;   // shear history written by this step
  const double4* k4;
  const double4* mx4;
  double4* D4;       // (D0, step of the break or 0)
  double k, ks, mu_d, mu_s;
  double truncated, beta1, beta2;
};
struct ForceLaunch {
  int count, nblocks, do2d;
  double period;  // circular x: particle 2 is imaged next to particle 1 (0: not circular)
  int ss;  // doubles per slot (6, or 9 with the contact point)
  int ublk;  // 1: block 0 runs the incremental update (queued steps), segments follow
  double dt;
  double step;  // this step's number (bond break stamps)
  double* slots;
  unsigned char* sact;  // per slot: 1 while it holds a nonzero share (zero slots are not read)
  const int2* slot_of;
  int* broken_list;  // this step's break list
  double* crit;      // per global bond index: breaking criterion of this step, or nullptr
  ForceSeg seg[ESP_MAX_FORCE_SEGS];
};
// TSubLattice::updateInteractions after bond breaks, for a queued step: the
// previous step's breaks are applied and the first dynamic group takes the
// new pairs of flagged particles; its segment in the force launch is useg.
struct UpdParams {
  BondSet bs;
  const int* broken_prev;  // the previous step's break list
  int4* evt;
  long long step;
  GridParams g;
  DynAppend da;
  const int* cell;
  const int* cell_start;
  const int* cell_items;
  UpdState* st;
  int* flist;
  int* xcnt;
  int has_group;
  int useg;
};
// Inputs only the extended force launch reads (k_forces_ext: rigid=True
// RotFriction groups, the pair-event log); the common launch takes none of
// them, so its code is the same with or without these features.
struct ForceExt {
  int* app_step;  // step each appendix pair is created in (tracked group), or nullptr
};
// blocks a segment of n interactions takes (dynamic groups: n + ESP_APP_CAP)
int force_blocks(int n);
// up: the incremental update of a queued step (fl.ublk = 1), or nullptr;
// ev_a/ev_b: events the launch stamps at the kernel's start and end; ext:
// the extended launch's inputs (nullptr: none)
bool launch_forces(hipStream_t s, StepCtl ctl, PView p, const ForceLaunch& fl,
                   const UpdParams* up, hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr,
                   const ForceExt* ext = nullptr);
// BrittleBeamSC groups: the stress test after the step's integration, one
// segment of blocks per group
struct ScLaunch {
  int count, nblocks;
  double period;  // circular x period (0: not circular)
  int blk0[ESP_MAX_BOND_GROUPS], n[ESP_MAX_BOND_GROUPS], bond_off[ESP_MAX_BOND_GROUPS];
  const int2* ij[ESP_MAX_BOND_GROUPS];
  double4* D4[ESP_MAX_BOND_GROUPS];
  double sigma3[ESP_MAX_BOND_GROUPS], cohesion[ESP_MAX_BOND_GROUPS], tan_angle[ESP_MAX_BOND_GROUPS];
  const double* sig;
  int* broken_list;  // this step's break list
  double step;
};
void launch_sc_break(hipStream_t s, StepCtl ctl, PView p, const ScLaunch& sl);
// PIS_EB::update on the device: bonds of the previous step's break list die,
// their particles are flagged, events are logged.  erase_grid: a table build
// follows in this step; bonds with a particle outside that grid are skipped
// (the build erases them).
void launch_break_apply(hipStream_t s, StepCtl ctl, PView p, BondSet bs, const int* broken_list,
                        UpdState* st, int* flist, int4* evt, long long step,
                        const GridParams* erase_grid = nullptr);
void launch_hash_clear(hipStream_t s, HashRef h);
void launch_hash_insert(hipStream_t s, HashRef h, const long long* keys, int n);
void launch_bond_keys(hipStream_t s, int nb, const int2* ij, const int* id, long long* keys);
void launch_alive_flags(hipStream_t s, int nb, const int2* ij, int* keep);
bool launch_gather_integrate(hipStream_t s, StepCtl ctl, PView p, const GatherParams& gp,
                             hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr);
// wall forces [step][nwg][3] from the partials [step][block][nwg][3]
void launch_wall_finalize(hipStream_t s, const double* part, int nblocks, int nwg, int nsteps,
                          double* out);
void launch_clear_status(hipStream_t s, StepStatus* ring, int ring_size, long long first_step,
                         int count, int* halted);
int gather_blocks(int n);

// ---- rebuild / update kernels ----
void launch_cells_count(hipStream_t s, PView p, GridParams g, int* cell, int* cell_count,
                        int* n_out);
void launch_cells_fill(hipStream_t s, int n, const int* cell, const int* cell_start,
                       int* cell_fill, int* cell_items);
void launch_cells_sort(hipStream_t s, int ncell, const int* cell_start, int* cell_items);
// candidate pairs from each particle in slot order, oriented (lower id, higher
// id), keyed by the slot pair; flagged = getNewList
void launch_pairs(hipStream_t s, int mode_emit, int flagged, PView p, GridParams g, double alpha,
                  const int* cell, const int* cell_start, const int* cell_items, int* cnt,
                  const int* off, int* out_p, int* out_q, long long* out_key);
void launch_clear_flag_bit(hipStream_t s, PView p, UpdState* st);
// candidates of wall group w (any order): slots, flags, positions, initial
// positions; *cnt counts all, the first cap are stored
void launch_wall_cands(hipStream_t s, PView p, int w, const double4* init, int cap, int* cnt,
                       int* slot, unsigned* flg, double4* pos, double4* ipos);
// one candidate particle of a wall servo, as the host iterates it
struct WallCand {
  double4 pos;         // position and radius
  double ix, iy, iz;   // initial position (bonded walls, else 0)
  unsigned flags;
  unsigned pad;
};
static_assert(sizeof(WallCand) == 64, "one servo candidate is one 64-byte record");
// the candidates of wall group w in particle-list order: mark[o] = 1 when the
// particle of creation order o (slot ord_slot[o], -1 erased) has the group's
// membership bit; pack writes the marked ones at out[off[o]] (off: the
// exclusive scan of mark); init null: no initial positions
void launch_wall_mark_ordered(hipStream_t s, int no, const int* ord_slot, const unsigned* flags,
                              int w, int* mark);
void launch_wall_pack_ordered(hipStream_t s, int no, const int* ord_slot, const int* mark,
                              const int* off, PView p, const double4* init, WallCand* out);
// a[0..n) = v (a kernel, not a memset: safe under stream capture)
void launch_fill_int(hipStream_t s, int n, int* a, int v);
// erasure of particles outside the padded grid (NeighborTable::build):
// keep[i] = cell[i] >= 0; nslot[i] = new slot or -1; stable compaction of a
// per-particle field of `width` elements; interaction pairs remapped to the new
// slots (keep: both particles survive); a dynamic group's kept pairs with keys
void launch_keep_cells(hipStream_t s, int n, const int* cell, int* keep);
void launch_new_slot(hipStream_t s, int n, const int* keep, const int* off, int* nslot);
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const double4* src, double4* dst);
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const double* src, double* dst);
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const int* src, int* dst);
void launch_compact_field(hipStream_t s, int n, int width, const int* keep, const int* off,
                          const unsigned* src, unsigned* dst);
void launch_remap_pairs(hipStream_t s, int nc, const int2* ij, const int* nslot, int2* out,
                        int* keep);
void launch_pack_dyn(hipStream_t s, int nc, const int* keep, const int* off, const int2* ij,
                     const double4* ff, int2* oij, long long* okey, double4* off_);
void launch_ed_mark(hipStream_t s, int nc, const int* cp, const int* cq, const long long* ckey,
                    ExclSet ex, const long long* okey, int nold, const int2* oij, PView p,
                    double period, int* keep, int* src);
// circular x: particles beyond an x end move to their image (position, old
// position and, when kept, initial position)
void launch_wrap_x(hipStream_t s, PView p, double x0, double period, double4* init);
void launch_ne_mark(hipStream_t s, int nc, const int* cp, const int* cq, const int* id, ExclSet ex,
                    int* keep);
void launch_pair_emit(hipStream_t s, int nc, const int* keep, const int* off, const int* cp,
                      const int* cq, const long long* ckey, const int* src, const double4* off_ff,
                      int2* nij, long long* nkey, double4* nff);
void launch_filter_new(hipStream_t s, int nc, const int* cp, const int* cq, const int* id,
                       const long long* ckey, ExclSet ex, HashRef present, int check_present,
                       int* keep);
void launch_merge(hipStream_t s, int nold, const int2* oij, const long long* okey,
                  const double4* off_, int nnew, const int2* nij_in, const long long* nkey_in,
                  const double4* nff_in, int2* ij_out, long long* key_out, double4* ff_out);
void launch_pack_pairs(hipStream_t s, int n, const int* keep, const int* off, const int* cp,
                       const int* cq, const long long* ckey, int2* ij, long long* key);
void launch_compact_bonds(hipStream_t s, int nb, const int* keep, const int* off,
                          const int2* ij, const double4* k4, const double4* mx4,
                          const double4* D4, const int* rank, int2* ij2, double4* k42,
                          double4* mx42, double4* D42, int* rank2);
void launch_wall_update(hipStream_t s, PView p, GridParams g, int nwg, const WallParam* walls);
void launch_inc_count(hipStream_t s, int nc, const int2* ij, int* inc_cnt);
// rank: the interactions' positions in the group's reference order (nullptr:
// storage order); grp: the group's position in the summation order (< 32)
void launch_inc_fill(hipStream_t s, int nc, const int2* ij, int rec_off, int has_m, const int* rank,
                     int grp, const int* inc_off, int* inc_fill, unsigned long long* ent64);
void launch_inc_sort(hipStream_t s, int n, const int* inc_off, unsigned long long* ent64,
                     int* inc_ent);
// slot_of[record] = (CSR position on its first particle, on its second)
void launch_slot_of(hipStream_t s, int total, const int* inc_ent, int2* slot_of);

// exclusive scan of n ints into out (out[n] = total); tmp sized by scan_tmp_size
size_t scan_tmp_size(int n);
void launch_exclusive_scan(hipStream_t s, const int* in, int* out, int n, int* tmp);
