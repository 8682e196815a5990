/*
 * esp.h - C-ABI of the MI355X per-timestep DEM particle engine.
 *
 * One handle (esp_engine) is one subdomain: one device-resident particle set
 * with its neighbour table, interaction groups, walls and body forces.  Every
 * entry point is extern "C", takes plain pointers and sizes, returns an int
 * status (ESP_OK == 0) and never throws; the message of the last failure is
 * available from esp_last_error().  Calls on one handle are not re-entrant and
 * must come from one host thread; all device work runs on the handle's stream.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the ESyS-Particle 2.3.5 source tree).  In the reference these calls cross
 * the Boost.Python -> CLatticeMaster -> MPI -> TSubLattice chain; here the
 * Python facade (esys-particle_amd/lsm.py) calls them directly through ctypes.
 *
 * Particle arrays passed in or out are structure-of-arrays in the caller's
 * order for inputs; outputs are ordered by ascending particle id.
 */
#ifndef ESP_H
#define ESP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ESP_OK 0
#define ESP_ERR_ARG -1      /* bad argument / unknown name */
#define ESP_ERR_STATE -2    /* call not allowed in the current state */
#define ESP_ERR_DEVICE -3   /* HIP runtime failure */
#define ESP_ERR_CAPACITY -4 /* size above a documented limit */
#define ESP_ERR_NUMERIC -5  /* non-finite state detected */

#define ESP_PTYPE_NROT 0 /* "NRotSphere" -> CParticle   (LsmMpiPy.cpp:227-256) */
#define ESP_PTYPE_ROT 1  /* "RotSphere"  -> CRotParticle */

/* particle fields for esp_get_particle_data */
#define ESP_FIELD_POS 0     /* 3 doubles per particle */
#define ESP_FIELD_VEL 1     /* 3 */
#define ESP_FIELD_ANGVEL 2  /* 3 (zeros for NRot) */
#define ESP_FIELD_QUAT 3    /* 4: scalar, x, y, z (identity for NRot) */
#define ESP_FIELD_FORCE 4   /* 3: total force of the last step */
#define ESP_FIELD_MOMENT 5  /* 3: total moment of the last step */
#define ESP_FIELD_RAD 6     /* 1 */
#define ESP_FIELD_MASS 7    /* 1 */
#define ESP_FIELD_OLDPOS 8  /* 3: position at the last neighbour search */

/* Lattice configuration.  Mirrors LsmMpi.initNeighbourSearch(particleType,
 * gridSpacing, verletDist) (LsmMpiPy.cpp:227-256, LatticeMaster.cpp:528-553),
 * setSpatialDomain (TSubLattice::initNeighborTable, SubLattice.hpp:455-478),
 * setTimeStepSize (LatticeMaster.cpp:574) and the 2-D switch
 * (CParticle::setDo2dCalculations, Particle.h:185). */
typedef struct esp_config {
  int particle_type;     /* ESP_PTYPE_* */
  double grid_spacing;   /* neighbour-table cell size, >= 2*r_max + verlet_dist */
  double verlet_dist;    /* alpha: search cutoff added to r_i + r_j */
  double domain_min[3];  /* spatial domain (grown to a multiple of grid_spacing) */
  double domain_max[3];
  double dt;             /* time step size */
  int do_2d;             /* 1: particles move in the x-y plane only */
  int device;            /* HIP device ordinal */
  int steps_per_batch;   /* steps queued per host synchronisation (>=1) */
  int circular_x;        /* 1: periodic in x (setSpatialDomain(bbox, circDimList=[1,0,0]),
                            LsmMpiPy.cpp:1093-1100 -> TSubLattice::initNeighborTable(min, max,
                            circ), SubLattice.hpp:495-564): the cell size is adjusted so that
                            an integer number of cells spans x, particles crossing an x end
                            re-enter at the other (wrapped at each neighbour-table rebuild,
                            as CParticle::setCircular, Particle.cpp:336-342) and every
                            interaction sees the image of its higher-id particle nearest to
                            its lower-id particle.  y and z cannot be circular. */
} esp_config;

/* RotBondPrms / BrittleBeamPrms after the Python-side mapping
 * (InteractionParamsPy.cpp:414-667, RotBondedInteraction.cpp:40-94).
 * beta1/beta2 are fixed at 1 by the reference's worker (RotSubLattice.hpp:123-166)
 * and are exposed here for completeness. */
typedef struct esp_rot_bond_prms {
  double kr, ks, kt, kb;                 /* normalK, shearK, torsionK, bendingK */
  double max_n, max_s, max_t, max_b;     /* normal/shear/torsion/bending break values */
  int tag;                               /* connection tag the group binds */
  int scaling;                           /* scale with particle size (default 1) */
  int meanR_scaling;                     /* mean radius (1) or min radius (0) */
  double truncated;                      /* tensile truncation (default 1) */
  double beta1, beta2;                   /* bending / torsion weights (default 1) */
} esp_rot_bond_prms;

/* BrittleBeamSCPrms(name, youngsModulus, poissonsRatio, cohesion, frictionAngle,
 * tensileCutoff, tag) after the Python-side mapping: kr = kb = pi E,
 * ks = kt = pi E / (2 (1 + nu)) (InteractionParamsPy.cpp:680-693,
 * ARotBondedInteraction.cpp:94-110, BrittleBeamSC.cpp:29-56). */
typedef struct esp_brittle_beam_sc_prms {
  double kr, ks, kt, kb;   /* stiffnesses before size scaling */
  double cohesion;         /* Mohr-Coulomb cohesion */
  double friction_angle;   /* degrees */
  double tension_cutoff;   /* tensile strength */
  int tag;                 /* connection tag the group binds */
} esp_brittle_beam_sc_prms;

/* RotFrictionPrms (InteractionParamsPy.cpp:772-833, RotFricInteraction.cpp:34-57). */
typedef struct esp_rot_friction_prms {
  double k_n;        /* normalK */
  double mu_d;       /* dynamicMu */
  double mu_s;       /* staticMu */
  double k_s;        /* shearK */
  int scaling;       /* default 1 */
  int rigid;         /* default 0; 1: calcRigidBodyForces (RotFricInteraction.cpp:237-343) */
  int meanR_scaling; /* default 1 */
} esp_rot_friction_prms;

/* Counters of one engine.  Interaction counts are the live totals. */
typedef struct esp_stats {
  long long time_step;       /* steps completed (CLatticeMaster::m_t) */
  long long num_rebuilds;    /* neighbour-table rebuilds (ppa timestamp) */
  long long num_particles;
  long long num_bonds;       /* live bonds over all bonded groups */
  long long num_dynamic;     /* interactions over all dynamic groups */
  long long num_candidates;  /* pairs of the last full neighbour list */
  long long num_broken;      /* bond-break events so far */
  long long num_host_phases; /* steps that needed host-side bookkeeping */
} esp_stats;

typedef struct esp_engine esp_engine;

/* ---- lifecycle ---------------------------------------------------------- */
/* LsmMpi(...) + initNeighbourSearch + setSpatialDomain + setTimeStepSize
 * (LsmMpiPy.cpp:196-256, 2991-3010; SubLattice.hpp:455-478). */
int esp_create(const esp_config* cfg, esp_engine** out);
void esp_destroy(esp_engine* e);
/* Message of the last failure on this handle ("" when none). */
const char* esp_last_error(const esp_engine* e);
/* Message of the last esp_create failure (no handle exists then). */
const char* esp_create_error(void);
/* setTimeStepSize (LatticeMaster.cpp:574, SubLattice.hpp:2163-2210). */
int esp_set_time_step_size(esp_engine* e, double dt);

/* ---- particles ---------------------------------------------------------- */
/* createParticles (LsmMpiPy.cpp:383-426) -> NeighborTable::insert
 * (ntable.hpp:231-255): particles outside the padded grid are dropped.
 * mass may be NULL (mass = r^3, SimpleParticleData.hpp:159-162),
 * vel may be NULL (zero).  tag may be NULL (0). */
int esp_add_particles(esp_engine* e, int n, const int* id, const int* tag,
                      const double* pos, const double* rad, const double* mass,
                      const double* vel);
/* setParticleDensity (SubLattice.hpp:2990-3001, Particle.cpp:108-122,
 * RotParticle.cpp:160-177). */
int esp_set_particle_density(esp_engine* e, int tag, int mask, double rho);
/* setParticleNonDynamic / setParticleNonRotational (Particle.h:137,
 * RotParticle.h:204-206): applied to particles with the given tag. */
int esp_set_particle_non_dynamic(esp_engine* e, int tag);
int esp_set_particle_non_rotational(esp_engine* e, int tag);
/* setParticleVelocity / setParticleAngularVelocity (LsmMpiPy.cpp:1234-1242 ->
 * CLatticeMaster::setParticleVel / setParticleAngVel, LatticeMaster.cpp:1499):
 * the particle with the given id, before or between time steps; the angular
 * form needs rotational particles. */
int esp_set_particle_vel(esp_engine* e, int id, const double v[3]);
int esp_set_particle_ang_vel(esp_engine* e, int id, const double w[3]);

/* ---- bonds -------------------------------------------------------------- */
/* createConnections (LsmMpiPy.cpp:284-288, SubLattice.hpp:600-620). */
int esp_add_connections(esp_engine* e, int n, const int* id1, const int* id2,
                        const int* tag);
/* createInteractionGroup(RotBondPrms|BrittleBeamPrms): TRotSubLattice::addRotBondedIG
 * (RotSubLattice.hpp:116-186) + PIS_E::tryInsert (pi_storage_e.hpp:216-250).
 * per_bond_max (NULL or 4*n_per_bond doubles: max_n, max_s, max_t, max_b before
 * size scaling) overrides the group strengths bond by bond, in the order of the
 * group's connections; this carries the per-bond strengths the reference can only
 * load from its restart format (RotBondedInteraction.cpp:406-428). */
int esp_add_rot_bond_group(esp_engine* e, const char* name,
                           const esp_rot_bond_prms* prms,
                           const double* per_bond_max, int n_per_bond);

/* createInteractionGroup(BrittleBeamSCPrms): LsmMpiPy::createBrittleBeamSCInteractGrp
 * (LsmMpiPy.cpp:645-649) -> TRotSubLattice::addBrittleBeamSCIG
 * (RotSubLattice.hpp:188-225): ARotBondedInteraction bonds from the connections
 * with the tag; a bond breaks on the mean stress of its two particles
 * (BrittleBeamSCInteraction::broken, BrittleBeamSC.cpp:103-148).  With such a
 * group the engine accumulates every particle's stress tensor
 * (CParticle::applyForce, Particle.cpp:391-400). */
int esp_add_brittle_beam_sc_group(esp_engine* e, const char* name,
                                  const esp_brittle_beam_sc_prms* prms);

/* ---- dynamic pair groups ------------------------------------------------ */
/* createInteractionGroup(NRotElasticPrms): ParallelInteractionStorage_NE
 * <T,CElasticInteraction> (SubLattice.hpp:1038-1054, ElasticInteraction.cpp:25-71). */
int esp_add_nrot_elastic_group(esp_engine* e, const char* name, double k, int scaling);
/* createInteractionGroup(RotElasticPrms): PIS_NE<CRotParticle,CRotElasticInteraction>
 * (RotElasticInteraction.cpp:40-88). */
int esp_add_rot_elastic_group(esp_engine* e, const char* name, double kr, int scaling);
/* createInteractionGroup(RotFrictionPrms): PIS_ED<CRotParticle,CRotFrictionInteraction>
 * (SubLattice.hpp:1202-1227, RotFricInteraction.cpp:103-231). */
int esp_add_rot_friction_group(esp_engine* e, const char* name,
                               const esp_rot_friction_prms* prms);
/* createExclusion(bonded, dynamic): TSubLattice::setExIG (SubLattice.hpp:1980-2009)
 * -> PIS_ED::addExIG / PIS_NE::addExIG (pi_storage_ed.hpp:20-40). */
int esp_add_exclusion(esp_engine* e, const char* excluding, const char* excluded);
/* createInteractionGroupTagged(prms, tag1, mask1, tag2, mask2) (LsmMpiPy.cpp:
 * 750-835 -> PIS_ED_T / PIS_NE_T, pi_storage_ed_t.hpp, pi_storage_ne_t.hpp):
 * the dynamic group just created takes only pairs whose smaller tag t1 and
 * larger tag t2 satisfy (t1 & m1) == (T1 & m1) and (t2 & m2) == (T2 & m2),
 * (T1, m1), (T2, m2) the given pairs sorted by tag.  Before the first step. */
int esp_set_interaction_group_tags(esp_engine* e, const char* name, int tag1, int mask1,
                                   int tag2, int mask2);

/* ---- walls and body forces ---------------------------------------------- */
/* createWall (LsmMpiPy.cpp:1145-1151, Wall.cpp:35-43). */
int esp_add_wall(esp_engine* e, const char* name, const double origin[3],
                 const double normal[3]);
/* createInteractionGroup(NRotElasticWallPrms): CEWallInteractionGroup
 * (EWallInteractionGroup.hpp:42-82, EWallInteraction.hpp:26-78). */
int esp_add_elastic_wall_group(esp_engine* e, const char* name, const char* wall,
                               double k);
/* createInteractionGroup(NRotBondedWallPrms(name, wallName, normalK,
 * particleTag, tagMask=-1)): CBWallInteractionGroup (BWallInteractionGroup.hpp,
 * BWallInteraction.hpp:22-52): of the wall's candidate particles those with
 * (tag & mask) == (particleTag & mask) are bonded to the wall (force
 * -(particle displacement since creation - wall displacement) k pi r), the
 * others interact elastically as with esp_add_elastic_wall_group. */
int esp_add_bonded_wall_group(esp_engine* e, const char* name, const char* wall, double k,
                              int tag, int mask);
/* createInteractionGroupTagged(NRotElasticWallPrms, tag, mask)
 * (LsmMpiPy.cpp:1203-1206 -> CTaggedEWallInteractionGroup,
 * TaggedEWallInteractionGroup.hpp:28-52): the elastic wall group just created
 * acts only on candidates with (tag & mask) == (particleTag & mask). */
int esp_set_wall_group_tags(esp_engine* e, const char* name, int tag, int mask);
/* createInteractionGroup(GravityPrms): BodyForceGroup (BodyForceGroup.hpp:41-65). */
int esp_add_gravity(esp_engine* e, const char* name, const double g[3]);
/* createInteractionGroup(DampingPrms / LinDampingPrms / RotDampingPrms /
 * LocalDampingPrms / RotLocalDampingPrms) -> TSubLattice::doAddDamping
 * (SubLattice.hpp:1747-1783, RotSubLattice.hpp:78-112).  Damping groups act
 * after every other force (SubLattice.hpp:2143-2147) in the order of their map
 * keys: the type name for ESP_DAMPING ("Damping") and ESP_ROT_DAMPING
 * ("RotDamping"), the group name for the local forms; a second group with an
 * existing key is ignored, as the reference's map insert ignores it.  The
 * rotational forms need rotational particles.  max_iterations is used by
 * ESP_DAMPING and ESP_ROT_DAMPING only (the reference velocity is zero). */
#define ESP_DAMPING 0           /* CDamping (Damping.hpp:87-107) */
#define ESP_ROT_DAMPING 1       /* CRotDamping (RotDamping.hpp:58-84) */
#define ESP_LOCAL_DAMPING 2     /* CLocalDamping (LocalDamping.hpp:82-117) */
#define ESP_ROT_LOCAL_DAMPING 3 /* CRotLocalDamping (RotLocalDamping.hpp:82-117) */
int esp_add_damping_group(esp_engine* e, int type, const char* name, double viscosity,
                          int max_iterations);
/* moveWallBy (LsmMpiPy.cpp:1161, SubLattice.hpp:2878-2893): immediate. */
int esp_move_wall_by(esp_engine* e, const char* name, const double d[3]);
/* applyForceToWall(interactionName, force) (LsmMpiPy.cpp:1208 ->
 * TSubLattice::applyForceToWall, SubLattice.hpp:2935-2950): the wall of the
 * wall interaction group `name` moves along the force so that the group's wall
 * force balances it: CEWallInteractionGroup::applyForce (elastic groups,
 * EWallInteractionGroup.hpp:93-164: at most 50 iterations of a stiffness-weighted
 * step until the relative force error is <= 1e-3) or CBWallInteractionGroup::
 * applyForce (bonded groups, BWallInteractionGroup.hpp:71-145: at most 10, until
 * the step is <= 1e-5).  Immediate, like moveWallBy; a servo-loaded wall calls
 * it from a pre-time-step runnable.  ESP_ERR_NUMERIC for a bonded group without
 * stiffness (the reference would move its wall by an infinite distance). */
int esp_apply_force_to_wall(esp_engine* e, const char* name, const double force[3]);
/* Checks the device bookkeeping the queued steps index with: per-particle
 * appendix slot counts, the flagged-particle list, the event count, every
 * stored pair and bond (slots in range, keys matching, force-record slots
 * inside their particles' summation runs), appendix counts.  ESP_OK, or
 * ESP_ERR_STATE with the first inconsistencies in esp_last_error.  No
 * counterpart in the reference: a consistency check for tests and debugging
 * (it synchronises and copies the bookkeeping to the host). */
int esp_check_state(esp_engine* e);
/* Maximum candidate particles of one wall group for esp_apply_force_to_wall_noalloc. */
#define ESP_NOALLOC_MAX_WALL_CANDIDATES 262144
/* esp_apply_force_to_wall without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (a servo calling it every time
 * step inside a tight loop, code under stream capture, harnesses that forbid
 * hipMalloc): esp_apply_force_to_wall allocates and frees its candidate
 * staging on every call; this variant stages them in a fixed file-scope device
 * array (about 19 MB per GPU, held until the process ends) and file-scope host
 * arrays.  Same result bit for bit.  ESP_ERR_CAPACITY when the group has more
 * than ESP_NOALLOC_MAX_WALL_CANDIDATES candidates.  Calls that overlap on one
 * GPU, or run concurrently in one process, race on the shared staging. */
int esp_apply_force_to_wall_noalloc(esp_engine* e, const char* name, const double force[3]);
/* A pre-time-step runnable that calls moveWallBy(name, d) at the start of every
 * step (compbench.py:20-35 pattern, LatticeMaster.cpp:2647).  d = 0 removes it. */
int esp_set_wall_step_motion(esp_engine* e, const char* name, const double d[3]);

/* ---- stepping ----------------------------------------------------------- */
/* nsteps x runTimeStep (LsmMpiPy.cpp:1061-1064 -> CLatticeMaster::runOneStep,
 * LatticeMaster.cpp:2602-2725 -> TSubLattice::checkNeighbors / searchNeighbors /
 * updateInteractions / oneStep, SubLattice.hpp:2224-2536). */
int esp_run_time_steps(esp_engine* e, int nsteps);

/* ---- queries ------------------------------------------------------------ */
int esp_get_num_particles(esp_engine* e);
int esp_get_stats(esp_engine* e, esp_stats* out);
/* getTimeStep (LsmMpiPy.cpp:466-468 -> CLatticeMaster::getTimeStep,
 * LatticeMaster.h:226): the steps completed, answered from the host without
 * touching the device (esp_get_stats also folds the device's event logs and
 * reads the interaction counts back). */
int esp_get_time_step(esp_engine* e, long long* out);
/* Particle ids, ascending. */
int esp_get_particle_ids(esp_engine* e, int* out);
/* One ESP_FIELD_* for every particle, ordered by ascending id. */
int esp_get_particle_data(esp_engine* e, int field, double* out);
/* getWallPosition / getWallForce (LsmMpiPy.cpp:2767-2786, LatticeMaster.cpp:993-1019). */
int esp_get_wall_position(esp_engine* e, const char* name, double out[3]);
int esp_get_wall_force(esp_engine* e, const char* name, double out[3]);
/* Wall force after each of the last `n` steps (oldest first); returns the count. */
int esp_get_wall_force_history(esp_engine* e, const char* name, int n, double* out);
/* Number of interactions stored in a group (dynamic or bonded), -1 if unknown. */
int esp_get_interaction_count(esp_engine* e, const char* name);
/* Particle-id pairs of a group (id1 < id2), sorted by (id1, id2). */
int esp_get_interactions(esp_engine* e, const char* name, int* id1, int* id2);
/* Shear-spring history of a RotFriction group, same order as esp_get_interactions:
 * ffric 3 doubles per contact, flags bit0 = slipping, bit1 = touching. */
int esp_get_friction_state(esp_engine* e, const char* name, double* ffric, int* flags);
/* Pairs of the last full neighbour list (NeighborTable::getFullList,
 * ntable.hpp:590-621), sorted by (id1, id2); returns the count when ids are NULL. */
int esp_get_neighbour_pairs(esp_engine* e, int* id1, int* id2);
/* Restart checkpoint in the reference's text format: <prefix>_t=<step>_0.txt
 * (CheckPointInfo::write, CheckPointInfo.cpp:73-96) naming <prefix>_t=<step>_1.txt
 * (TSubLattice::saveCheckPointData, SubLattice.hpp:3599-3672: particles, bonded
 * groups, RotFriction groups, walls), `precision` significant digits (the
 * reference's RestartCheckPointPrms default: 12).  As the reference does before
 * a checkpoint (LatticeMaster.cpp:2697-2700), the neighbour table is rebuilt
 * first, so the next step updates every interaction group.  num_time_steps is
 * recorded in the info file (setNumTimeSteps of the saving script). */
int esp_save_checkpoint(esp_engine* e, const char* prefix, int num_time_steps, int precision);
/* LsmMpiPy::loadCheckPoint (LsmMpiPy.cpp:880-883) -> CLatticeMaster::loadCheckPointData
 * (LatticeMaster.cpp:780-820): before the first time step, with the script's
 * interaction groups and walls created and no particles yet: particles with their
 * state, bonds with their stored constants (matched to the bonded groups in name
 * order), RotFriction shear histories, wall positions; time steps continue from
 * the checkpoint's. */
int esp_load_checkpoint(esp_engine* e, const char* info_file);
/* Stress sums of the last step, 9 doubles per particle (row-major), ascending
 * id order: CParticle::m_sigma (Particle.cpp:391-400) before the division by
 * the particle volume.  ESP_ERR_STATE unless a BrittleBeamSC group exists. */
int esp_get_particle_stress(esp_engine* e, double* out);
/* Bond-break events (step, id1, id2) in (step, id1, id2) order; returns the count
 * when the pointers are NULL. */
int esp_get_bond_breaks(esp_engine* e, int* step, int* id1, int* id2);
/* For InteractionScalarFieldSaverPrms(bondGroup, "breaking_criterion", ...):
 * from the first time step on every bond stores
 * CRotBondedInteraction::getCriterion (RotBondedInteraction.cpp:221-227) of
 * the forces of its last step.  Before the first time step only. */
int esp_track_bond_criteria(esp_engine* e);
/* The live bonds of a bonded group in the group's list (connection) order
 * with their criterion of the last step (0 before the first step, as the
 * reference's zero-initialised cache); returns the count.  With NULL arrays
 * only the count is returned.  Needs esp_track_bond_criteria. */
int esp_get_bond_criteria(esp_engine* e, const char* name, int* id1, int* id2, double* crit);
/* Steps at which the neighbour table was rebuilt; returns the count when NULL. */
int esp_get_rebuild_steps(esp_engine* e, int* steps);
/* Log of the pairs a dynamic interaction group creates and removes
 * (PIS_NE::update / PIS_ED::update, pi_storage_ne.hpp:49-102,
 * pi_storage_ed.hpp:84-161, with the erasure of interactions whose particle
 * left the table, pi_storage_e.hpp:111-155).  Enabled per group before the
 * first time step.  From then on, after each step's interaction update, a pair
 * (id1, id2) whose number of stored interactions changed gives one event per
 * unit of change: kind +1 (created) or -1 (removed), step = that time step.
 * A pair dropped and created again in one update (a RotFriction pair that
 * stopped touching but is still a neighbour) is no event. */
int esp_track_pair_events(esp_engine* e, const char* name);
/* The events in (step, id1, id2) order, kind per event; returns the count when
 * the pointers are NULL.  ESP_ERR_STATE unless the group is tracked. */
int esp_get_pair_events(esp_engine* e, const char* name, int* step, int* id1, int* id2, int* kind);

/* ---- measurement (no counterpart in the reference) ----------------------- */
/* Kernel classes timed by esp_get_kernel_times. */
#define ESP_KCLASS_FORCES 0 /* force laws of every interaction group (+ break test), with a
                              queued step's bond-break bookkeeping in its block 0 */
#define ESP_KCLASS_GATHER 1 /* force gather + integration, per step */
#define ESP_KCLASS_COUNT 2
/* Enable (1) or disable (0) HIP-event timing of every per-step kernel launch on
 * the engine's stream: each launch stamps a (start, stop) event pair at the
 * kernel's own start and end (hipExtLaunchKernelGGL). */
int esp_set_kernel_timing(esp_engine* e, int enable);
/* Accumulated device time (ms) and launch count per kernel class, over the
 * steps that completed since timing was enabled; resets the accumulators. */
int esp_get_kernel_times(esp_engine* e, double* ms, long long* launches);

#ifdef __cplusplus
}
#endif

#endif /* ESP_H */
