/*
 * esp_diff.h - C-ABI of the differentiable, batched engine for micro-parameter
 * calibration (no counterpart in the reference: it adds gradients of a bonded
 * specimen's stress response with respect to its micro-parameters).
 *
 * Model class: RotSphere particles (dynamic, rotational, 3-D) of one or more
 * specimens in one particle set; one rotational bond group (RotBondPrms with
 * size scaling by the mean radius: kr = normalK, ks = shearK, kt, kb,
 * CRotBondedInteraction::broken with truncated = beta1 = beta2 = 1,
 * ARotBondedInteraction.cpp:307-448, RotBondedInteraction.cpp:122-219); one
 * RotFriction group over all unbonded neighbour pairs (mean-radius scaling,
 * static = dynamic coefficient, RotFricInteraction.cpp:103-231) excluding the
 * live bonds (createExclusion); elastic walls (NRotElasticWallPrms,
 * EWallInteraction.hpp:26-78) moved by a constant step at the start of every
 * step; optional gravity.  Step order, neighbour search (rebuild when a
 * particle moved more than verletDist/2), bond removal at the step after the
 * criterion held and the friction history rules are the reference's
 * (LatticeMaster.cpp:2602-2725, SubLattice.hpp:2224-2536).
 *
 * Parameters per specimen s: theta[4 s + 0..3] = (k_n, k_s, cohesion, mu):
 * k_n is the bonds' normalK and the contacts' normalK, k_s both shearK,
 * mu the contacts' static and dynamic friction coefficient, and each bond's
 * break values are cohesion times its relative strengths
 * (max_n, max_s, max_t, max_b before size scaling).
 *
 * Loss: the stress of each specimen - the normal force the measured wall
 * applies to the specimen's particles (the wall force CWall sums, along the
 * wall normal; positive in compression) over the specimen's area - is sampled
 * every `sample_every` steps (steps sample_every, 2 sample_every, ...);
 * loss_s = mean over the samples of (stress - target)^2.
 *
 * Gradients: esp_diff_backward differentiates every specimen's loss with
 * respect to its theta exactly along the branch the forward run took (contact
 * topology, slip states and bond breaks as recorded): a reverse sweep over
 * checkpointed segments of `checkpoint_every` steps, each recomputed forward
 * before its adjoint; the laws' local Jacobians come from their dual-number
 * evaluation (esp_difflaw.h).  The break criterion depends on cohesion only
 * through events, so d loss / d cohesion is zero on the branch.
 */
#ifndef ESP_DIFF_H
#define ESP_DIFF_H

#include "esp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct esp_diff_config {
  double grid_spacing;   /* neighbour-table cell size, >= 2 r_max + verlet_dist */
  double verlet_dist;    /* search cutoff alpha; rebuild after alpha/2 of displacement */
  double domain_min[3];  /* spatial domain (grown to whole cells, one padding cell) */
  double domain_max[3];
  double dt;
  int device;
  int checkpoint_every;  /* steps per checkpointed segment of the adjoint (>= 1) */
  int block;             /* threads per block of the per-step kernels (64..512, multiple of 64) */
} esp_diff_config;

typedef struct esp_diff esp_diff;

int esp_diff_create(const esp_diff_config* cfg, esp_diff** out);
void esp_diff_destroy(esp_diff* d);
const char* esp_diff_last_error(const esp_diff* d);
const char* esp_diff_create_error(void);

/* Particles 0..n-1 (the index is the particle id; lower index = lower id),
 * spec[i] the specimen, non-decreasing (each specimen's particles contiguous);
 * pos 3n, rad n, mass n (moment of inertia 0.4 m r^2), vel 3n or NULL. */
int esp_diff_set_particles(esp_diff* d, int n, int nspec, const int* spec, const double* pos,
                           const double* rad, const double* mass, const double* vel);
/* Bonds between particles i[k] < j[k] of the same specimen, with relative
 * break values rel[4 k + 0..3] (max_n, max_s, max_t, max_b per unit cohesion,
 * before size scaling); kt, kb per specimen (torsionK, bendingK, before
 * scaling) in tk[s], bk[s]. */
int esp_diff_set_bonds(esp_diff* d, int nb, const int* i, const int* j, const double* rel,
                       const double* tk, const double* bk);
/* An elastic wall (origin, unit normal, stiffness k before the pi r scaling)
 * moved by `step` at the start of every step; measure = 1 for the wall whose
 * force gives the stress (one wall), area[s] the specimens' loaded areas. */
int esp_diff_add_wall(esp_diff* d, const double origin[3], const double normal[3], double k,
                      const double step[3], int measure, const double* area);
int esp_diff_set_gravity(esp_diff* d, const double g[3]);
/* Steps of one run and the sampling interval of the stress. */
int esp_diff_set_run(esp_diff* d, int nsteps, int sample_every);

/* One forward run from the initial state with theta [4 nspec] against the
 * target stresses [nspec * nsamples]: the sampled stresses (may be NULL) and
 * the loss per specimen.  Keeps the checkpoints for esp_diff_backward. */
int esp_diff_forward(esp_diff* d, const double* theta, const double* target, double* stress,
                     double* loss);
/* d loss_s / d theta_s [4 nspec] of the last forward run. */
int esp_diff_backward(esp_diff* d, double* grad);

/* State after the last forward run: ESP_FIELD_POS / VEL / ANGVEL / QUAT
 * (particle index order). */
int esp_diff_get_particle_data(esp_diff* d, int field, double* out);
/* Bond breaks of the last forward run: (step at which the bond left the
 * group, i, j) in step then bond order; returns the count (NULL: count only). */
int esp_diff_get_bond_breaks(esp_diff* d, int* step, int* i, int* j);
/* Steps at which the neighbour table was rebuilt in the last forward run. */
int esp_diff_get_rebuild_steps(esp_diff* d, int* steps);
/* esp_diff_forward along a parameter direction: also carries the tangent of
 * the state (forward mode, on the branch the run takes) and returns
 * dloss[s] = d loss_s / d eps at theta + eps * dtheta (dtheta [nspec][4]).
 * An independent derivative of the same forward run, for checking the
 * adjoint (grad . dtheta); no reference counterpart. */
int esp_diff_forward_tangent(esp_diff* d, const double* theta, const double* dtheta,
                             const double* target, double* stress, double* loss, double* dloss);
/* Compute (1) or skip (0, the default) the branch digests of later forward
 * runs.  Every wave of the contact and gather kernels adds its hash to one
 * counter per step, so tracking costs forward time.  No reference
 * counterpart: it serves the finite-difference checks of the adjoint. */
int esp_diff_track_branches(esp_diff* d, int enable);
/* Branch digest of each step of the last forward run (steps 1..nsteps into
 * out[0..nsteps-1]): a hash sum of every contact's touching and slipping state
 * and every particle's wall contacts.  Runs whose digests and bond breaks are
 * equal took the same branches, so their loss is differentiable between them.
 * Returns nsteps (out NULL: count only); ESP_ERR_STATE unless digests are
 * tracked (esp_diff_track_branches).  No reference counterpart: it serves the
 * finite-difference checks of the adjoint. */
int esp_diff_get_branch_digest(esp_diff* d, unsigned long long* out);
/* Device time (ms) of the last forward and backward runs. */
int esp_diff_get_times(esp_diff* d, double* forward_ms, double* backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* ESP_DIFF_H */
