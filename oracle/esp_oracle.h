/*
 * esp_oracle.h - CPU restatement of ESyS-Particle's per-timestep particle engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * It exports the same entry points as include/esp.h with the prefix eso_
 * instead of esp_ and the same argument meaning; see esp_oracle.c for the
 * reference file:line each part restates.
 */
#ifndef ESP_ORACLE_H
#define ESP_ORACLE_H

#include "../include/esp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eso_engine eso_engine;

int eso_create(const esp_config* cfg, eso_engine** out);
void eso_destroy(eso_engine* e);
const char* eso_last_error(const eso_engine* e);
const char* eso_create_error(void);
int eso_set_time_step_size(eso_engine* e, double dt);
int eso_add_particles(eso_engine* e, int n, const int* id, const int* tag,
                      const double* pos, const double* rad, const double* mass,
                      const double* vel);
int eso_set_particle_density(eso_engine* e, int tag, int mask, double rho);
int eso_set_particle_non_dynamic(eso_engine* e, int tag);
int eso_set_particle_non_rotational(eso_engine* e, int tag);
int eso_add_connections(eso_engine* e, int n, const int* id1, const int* id2,
                        const int* tag);
int eso_add_rot_bond_group(eso_engine* e, const char* name,
                           const esp_rot_bond_prms* prms,
                           const double* per_bond_max, int n_per_bond);
int eso_add_nrot_elastic_group(eso_engine* e, const char* name, double k, int scaling);
int eso_add_rot_elastic_group(eso_engine* e, const char* name, double kr, int scaling);
int eso_add_rot_friction_group(eso_engine* e, const char* name,
                               const esp_rot_friction_prms* prms);
int eso_add_exclusion(eso_engine* e, const char* excluding, const char* excluded);
int eso_add_wall(eso_engine* e, const char* name, const double origin[3],
                 const double normal[3]);
int eso_add_elastic_wall_group(eso_engine* e, const char* name, const char* wall,
                               double k);
int eso_add_gravity(eso_engine* e, const char* name, const double g[3]);
int eso_add_bonded_wall_group(eso_engine* e, const char* name, const char* wall, double k,
                              int tag, int mask);
int eso_set_wall_group_tags(eso_engine* e, const char* name, int tag, int mask);
int eso_add_damping_group(eso_engine* e, int type, const char* name, double viscosity,
                          int max_iterations);
int eso_set_particle_vel(eso_engine* e, int id, const double v[3]);
int eso_track_bond_criteria(eso_engine* e);
int eso_set_interaction_group_tags(eso_engine* e, const char* name, int tag1, int mask1, int tag2,
                                   int mask2);
int eso_get_bond_criteria(eso_engine* e, const char* name, int* id1, int* id2, double* crit);
int eso_set_particle_ang_vel(eso_engine* e, int id, const double w[3]);
int eso_move_wall_by(eso_engine* e, const char* name, const double d[3]);
int eso_set_wall_step_motion(eso_engine* e, const char* name, const double d[3]);
int eso_run_time_steps(eso_engine* e, int nsteps);
int eso_get_num_particles(eso_engine* e);
int eso_get_stats(eso_engine* e, esp_stats* out);
int eso_get_time_step(eso_engine* e, long long* out);
int eso_get_particle_ids(eso_engine* e, int* out);
int eso_get_particle_data(eso_engine* e, int field, double* out);
int eso_get_wall_position(eso_engine* e, const char* name, double out[3]);
int eso_get_wall_force(eso_engine* e, const char* name, double out[3]);
int eso_get_wall_force_history(eso_engine* e, const char* name, int n, double* out);
int eso_get_interaction_count(eso_engine* e, const char* name);
int eso_get_interactions(eso_engine* e, const char* name, int* id1, int* id2);
int eso_get_friction_state(eso_engine* e, const char* name, double* ffric, int* flags);
int eso_get_neighbour_pairs(eso_engine* e, int* id1, int* id2);
int eso_get_bond_breaks(eso_engine* e, int* step, int* id1, int* id2);
int eso_get_rebuild_steps(eso_engine* e, int* steps);
int eso_track_pair_events(eso_engine* e, const char* name);
int eso_get_pair_events(eso_engine* e, const char* name, int* step, int* id1, int* id2, int* kind);
int eso_set_kernel_timing(eso_engine* e, int enable);
int eso_get_kernel_times(eso_engine* e, double* ms, long long* launches);

/* Oracle-only extras used by the tests. */
/* Particle stress tensors (9 doubles, row-major, unscaled sigma of
 * CParticle::applyForce, Particle.cpp:391-400), ascending id order. */
int eso_get_particle_stress(eso_engine* e, double* out);
int eso_save_checkpoint(eso_engine* e, const char* prefix, int num_time_steps, int precision);
int eso_load_checkpoint(eso_engine* e, const char* info_file);
int eso_add_brittle_beam_sc_group(eso_engine* e, const char* name,
                                  const esp_brittle_beam_sc_prms* sc);
/* Particle-id pairs of a group in storage order, i.e. the std::list order of
 * the reference's interaction storage; returns the count. */
int eso_get_group_list_order(eso_engine* e, const char* name, int* id1, int* id2);

int eso_apply_force_to_wall(eso_engine* e, const char* name, const double F[3]);
/* Per particle (ascending id): sum of the norms of the forces and sum of the
 * norms of the moments applied in the last step (2 doubles each). */
int eso_get_force_scale(eso_engine* e, double* out);

/* Single-law entry points (no engine): inputs are the two particles' state,
 * outputs the forces/moments the law applies, in the law's own arithmetic. */
/* ARotBondedInteraction::calcForces (ARotBondedInteraction.cpp:307-448).
 * p = {pos1[3], quat1[4], rad1, pos2[3], quat2[4], rad2} (16 doubles)
 * b = {kr, ks, kt, kb, D[3]} (7 doubles)
 * out = {F1[3], M1[3], F2[3], M2[3], nForce, shForce, tMoment, bMoment} (16) */
void eso_law_rot_bond(const double* p, const double* b, double* out);
/* CRotFrictionInteraction::calcSimpleForces (RotFricInteraction.cpp:171-231).
 * p = {pos1[3], vel1[3], angvel1[3], rad1, pos2[3], vel2[3], angvel2[3], rad2} (20)
 * prm = {k, ks, mu_d, mu_s, dt} (already scaled), state = {Ffric[3], slipping}
 * (updated in place), out = {F1[3], M1[3], F2[3], M2[3], touching} (13) */
void eso_law_rot_friction(const double* p, const double* prm, double* state, double* out);
void eso_law_rot_friction_rigid(const double* p, const double* prm, double* state, double* out);
/* CElasticInteraction::calcForces (ElasticInteraction.cpp:54-71):
 * p = {pos1[3], rad1, pos2[3], rad2}, k scaled; out = {F1[3], F2[3], touching} */
void eso_law_nrot_elastic(const double* p, double k, double* out);
/* Scaled constants of a rotational bond (RotBondedInteraction.cpp:122-161):
 * in = {pos1[3], rad1, pos2[3], rad2}, prm = {kr,ks,kt,kb,max_n,max_s,max_t,max_b,
 * scaling, meanR_scaling, do_2d}; out = {kr,ks,kt,kb,max_n,max_s,max_t,max_b,D[3]} */
void eso_bond_constants(const double* in, const double* prm, double* out);

/* Math hooks over the oracle's own Vec3/Matrix3/Quaternion restatement; the
 * tests compare them with oracle/_ref/libref_math.so (the reference's compiled
 * Foundation classes).  Quaternions {scalar, x, y, z}, matrices row-major. */
void eso_math_quat_mul(const double* a, const double* b, double* o);
void eso_math_quat_inverse(const double* a, double* o);
void eso_math_quat_to_matrix(const double* a, double* o);
void eso_math_mat_vec(const double* m, const double* v, double* o);
void eso_math_mat_trans(const double* m, double* o);
void eso_math_vec_cross(const double* a, const double* b, double* o);
double eso_math_vec_dot(const double* a, const double* b);
double eso_math_vec_norm(const double* a);
void eso_math_quat_step(const double* q, const double* w, double dt, double* o);
void eso_math_rot_particle_step(double* st, double dt);

#ifdef __cplusplus
}
#endif

#endif
