/*
 * starflood.h -- drop-in C ABI of Starflood's force-and-integration path,
 * served by libstarflood_hip.so (C host code + HIP kernels for gfx950).
 *
 * The types below are byte-identical to the reference's headers under its
 * default configuration (FP32 `real`, ENABLE_TREE defined, ENABLE_GRID not
 * defined), because callers allocate sim_t on their own stack
 * (reference src/main.c:97) and index sim.mem through sim_find.
 *
 * What stays the same as the reference:
 *   - every prototype (names, argument order and meaning, return codes
 *     STARFLOOD_SUCCESS / STARFLOOD_FAILURE, "func() error: ..." on stderr);
 *   - the layout of sim.mem (enum sim_conf blocks of N reals each);
 *   - the host-authoritative contract by default: after a call returns, the
 *     host copy of sim.mem holds the results (SF_MEM_MIRROR, see
 *     starflood_ext.h for the device-resident alternative).
 *
 * What changes:
 *   - compile-time configuration (reference src/config.h) becomes run-time
 *     parameters set through sf_set_params() (starflood_ext.h), whose
 *     defaults are the config.h values;
 *   - ENABLE_SPH still selects the caller's view of enum sim_conf; the
 *     library is told the same choice through sf_params_t.enable_sph;
 *   - sim_t always carries tree_t (the reference's config.h:35 defines
 *     ENABLE_TREE); Barnes-Hut versus direct summation is chosen at run time;
 *   - tree_build errors are propagated by sim_solv instead of being ignored
 *     (reference src/simulation.c:312 drops the return code).
 *
 * SOLVER_VIS (reference config.h:32, defined there by default) selects the
 * prototypes of sim_solv, sim_step and solve_gravity_part_tree that also take
 * the tree debug overlay's render buffer, as the reference headers do.
 * libstarflood_hip.so exports the plain forms; libstarflood_hip_solvervis.so
 * is the same library built with SOLVER_VIS, exporting the overlay forms,
 * which accept and ignore the buffer (the overlay is not drawn).
 * STARFLOOD_DOUBLE_PRECISION is not supported: `real` is binary32.
 */
#ifndef STARFLOOD_H
#define STARFLOOD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define SF_RESTRICT __restrict__
extern "C" {
#else
#define SF_RESTRICT restrict
#endif

/* reference src/types.h:7-29 */
typedef int16_t i16;
typedef int32_t i32;
typedef int64_t i64;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef float f32;
typedef double f64;
typedef f32 real;
#ifndef STARFLOOD_NO_UINT_TYPEDEF
typedef u32 uint;
#endif

/* reference src/common.h:31-32 */
#define STARFLOOD_SUCCESS 0
#define STARFLOOD_FAILURE -1

/* reference src/tree.h:7-13 */
enum tree_conf {
	TREE_MAS_X,
	TREE_MAS_Y,
	TREE_MAS_Z,
	TREE_MAS_W,
	TREE_DOF
};

/* reference src/tree.h:16-25: 80 bytes with FP32 real */
typedef struct {
	u32 parent;
	u32 bodies;
	u32 branch[8];
	real bounds[6];
	real params[TREE_DOF];
} node_t;

/* reference src/tree.h:28-35 */
typedef struct {
	u32 max_nodes;
	u32 num_nodes;
	real bounds[6];
	node_t* node;
} tree_t;

/* reference src/simulation.h:17-51 */
enum sim_conf {
	_SIM_POS_0,
	_SIM_POS_1,
	_SIM_POS_2,
	_SIM_VEL_0,
	_SIM_VEL_1,
	_SIM_VEL_2,
	_SIM_ACC_0,
	_SIM_ACC_1,
	_SIM_ACC_2,
	SIM_MAS,
	SIM_POT,
	SIM_KEN,
	SIM_PEN,
#ifdef ENABLE_SPH
	SIM_RAD,
	SIM_RHO,
	SIM_PRS,
#endif
	SIM_DOF,
	SIM_POS = _SIM_POS_0,
	SIM_VEL = _SIM_VEL_0,
	SIM_ACC = _SIM_ACC_0
};

/* reference src/simulation.h:53-67 (ENABLE_TREE defined, ENABLE_GRID not) */
typedef struct {
	unsigned int step_number;
	unsigned int N;
	void* mem;
	tree_t tree;
} sim_t;

/* replaces reference src/simulation.h:70 (src/simulation.c:33-53) */
real* sim_find(const sim_t* SF_RESTRICT sim_ptr, enum sim_conf parameter_enum);

/* replaces reference src/simulation.h:73 (src/simulation.c:55-187).
   sim.mem is 4 MiB aligned and zeroed as in the reference; from 8 MiB up it
   is page-locked host memory (free it only through sim_free). */
int sim_init(sim_t* SF_RESTRICT sim_ptr, unsigned int N);

/* replaces reference src/simulation.h:76 (src/simulation.c:189-228) */
int sim_free(sim_t* SF_RESTRICT sim_ptr);

/* replaces reference src/simulation.h:79 (src/simulation_io.c:16-79) */
int sim_read_raw(sim_t* SF_RESTRICT sim_ptr, const char* SF_RESTRICT filename);

/* replaces reference src/simulation.h:85 (src/simulation_io.c:81-149) */
int sim_save_raw(const sim_t* SF_RESTRICT sim_ptr, const char* SF_RESTRICT filename);

#ifdef SOLVER_VIS
/* replaces reference src/simulation.h:92 (src/simulation.c:231-345); the
   render buffer (the tree debug overlay) is ignored */
int sim_solv(sim_t* SF_RESTRICT sim_ptr, f32* SF_RESTRICT render_buffer, unsigned int sizex, unsigned int sizey);

/* replaces reference src/simulation.h:95 (src/simulation.c:349-519) */
int sim_step(sim_t* SF_RESTRICT sim_ptr, f32* SF_RESTRICT render_buffer, unsigned int sizex, unsigned int sizey);

/* replaces reference src/gravity.h:15 (src/gravity.c:245-663) */
int solve_gravity_part_tree(const sim_t* SF_RESTRICT sim_ptr, f32* render_buffer, unsigned int sizex,
                            unsigned int sizey);
#else
/* replaces reference src/simulation.h:98 (src/simulation.c:231-345) */
int sim_solv(sim_t* SF_RESTRICT sim_ptr);

/* replaces reference src/simulation.h:101 (src/simulation.c:349-519) */
int sim_step(sim_t* SF_RESTRICT sim_ptr);

/* replaces reference src/gravity.h:17 (src/gravity.c:245-663) */
int solve_gravity_part_tree(const sim_t* SF_RESTRICT sim_ptr);
#endif

/* replaces reference src/gravity.h:22 (src/gravity.c:666-819) */
int solve_gravity_part_part(const sim_t* SF_RESTRICT sim_ptr);

/* replaces reference src/sph.h:6 (src/sph.c:14-283); runs only when the
   library was told enable_sph = 1 */
int solve_sph(const sim_t* SF_RESTRICT sim_ptr);

/* replaces reference src/tree.h:38 (src/tree.c:14-55) */
int tree_init(tree_t* SF_RESTRICT tree_ptr, unsigned int max_nodes);

/* replaces reference src/tree.h:41 (src/tree.c:57-72) */
int tree_free(tree_t* SF_RESTRICT tree_ptr);

/* replaces reference src/tree.h:44 (src/tree.c:74-318) */
int tree_build(tree_t* SF_RESTRICT tree_ptr, unsigned int num_parts, const real* SF_RESTRICT pos, const real* SF_RESTRICT mas);

/* replaces reference src/initcond.h:7 (src/initcond.c:11-332) */
int initcond_generate(sim_t* SF_RESTRICT sim_ptr);

#ifdef __cplusplus
}
#endif

#endif /* STARFLOOD_H */
