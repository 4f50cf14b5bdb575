/*
 * sf_host.h -- solver entry points shared by the host translation units.
 * Each works on device state only (no host transfers, no synchronisation);
 * the exported reference-signature functions wrap them with the mirror-mode
 * transfers.
 */
#ifndef SF_HOST_H
#define SF_HOST_H

#include "../sf_internal.h"

/* gravity.c: PP gravity from d_mem pos/mas into d_mem acc/pot */
int sf_solve_gravity_pp(sf_ctx* c);
/* gravity.c: j ranges of the fast PP kernel (and of the SPH all-pairs and
   VJP passes) for n_pad particles, the bytes of slice partials the fast PP
   kernel writes, and the fast PP force on caller arrays (body: n_pad float4,
   part: part_bytes >= sf_pp_part_bytes(n_pad), pot may be NULL) */
unsigned int sf_pp_slice_count(unsigned int n_pad);
size_t sf_pp_part_bytes(unsigned int n_pad);
int sf_pp_force_arrays(const sf_params_t* p, unsigned int N, unsigned int n_pad, const float* pos, const float* mas,
                       float* body, float* part, size_t part_bytes, float* acc, float* pot, hipStream_t s);
/* tree.c: bounds as src/simulation.c:245-301, build, then walk into acc/pot.
   prev_acc: acc holds the previous force evaluation (sim_step), which the
   LBVH relative opening test reads; otherwise it first walks without it. */
int sf_solve_gravity_tree(sf_ctx* c, tree_t* tree, unsigned int step_number, int prev_acc);
/* tree.c: after an LBVH solve on stream s, fail if the walk met a malformed
   tree (synchronises s; no-op in reference mode) */
int sf_tree_check(tree_t* tree, hipStream_t s, const char* fn);
/* sph.c: density, EOS and pressure force, added into acc.  For the cubic
   model sf_sph_prepare packs the particle records first, while acc still
   holds the previous forces: velocities and u are predicted pred_dt ahead. */
int sf_sph_prepare(sf_ctx* c, float pred_dt);
int sf_solve_sph(sf_ctx* c);
/* simulation.c: sim_solv and sim_step, whichever prototype form the library
   exports (SOLVER_VIS or not) */
int sf_sim_solv(sim_t* SF_RESTRICT sim_ptr);
int sf_sim_step(sim_t* SF_RESTRICT sim_ptr);
/* simulation.c: the solver sequence of sim_solv on device state
   (sph_pred_dt: see sf_sph_prepare) */
int sf_solve_all(sf_ctx* c, tree_t* tree, unsigned int step_number, float sph_pred_dt, int prev_acc);

/* lbvh.c: SF_TREE_LBVH build (allocates *st on first use) and walk */
typedef struct sf_lbvh sf_lbvh;
int sf_lbvh_build(sf_lbvh** st, unsigned int N, const float* pos, const float* mas, int quadrupole, hipStream_t s);
int sf_lbvh_walk(sf_lbvh* st, unsigned int N, const float* pos, const float* mas, const sf_params_t* p, int prev_acc,
                 float* acc, float* pot, hipStream_t s);
void sf_lbvh_destroy(sf_lbvh* st);
/* particles of the LBVH last built (0: none, or no build yet) */
unsigned int sf_lbvh_size(const sf_lbvh* st);
/* walk wavefronts that met a malformed tree since the last call (cleared;
   synchronises s) */
int sf_lbvh_guard(sf_lbvh* st, hipStream_t s, unsigned long long* tripped);
/* synchronises s, then copies out what is asked for (NULL: skip) */
int sf_lbvh_export_impl(sf_lbvh* st, hipStream_t s, unsigned long long* keys, unsigned int* order,
                        unsigned long long counts[4], float* node_cm, int* node_lr_range, float* node_cell);

/* initcond.c: after a host-side write of sim.mem, refresh the device copy
   in SF_MEM_RESIDENT mode */
int sf_after_host_write(const sim_t* sim_ptr, const char* func);

#endif
