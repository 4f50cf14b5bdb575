/*
 * starflood_ext.h -- additive entry points of libstarflood_hip.so that have
 * no counterpart symbol in the reference.
 *
 * The reference configures everything at compile time (src/config.h); here
 * the same knobs are run-time parameters whose defaults are the config.h
 * values.  The other entry points exist because the device is a separate
 * memory: explicit host<->device synchronisation for the device-resident
 * mode, device pointers for zero-copy interop (PyTorch), FP64 statistics,
 * kernel timing, and the adjoint of the force and of the KDK step.
 */
#ifndef STARFLOOD_EXT_H
#define STARFLOOD_EXT_H

#include "starflood.h"
#include "starflood_vis.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic policy of the force, KDK and statistics kernels. */
enum sf_precision {
	/* Reference operation order in FP32: IEEE divide and sqrt, no FMA, per
	   pair Kahan, sequential j order per particle, FP32 drift, sequential
	   FP64 Kahan energy sums (reference src/gravity.c:693-812,
	   src/simulation.c:381-497).  Bit-identical to the reference CPU path. */
	SF_PRECISION_EXACT = 0,
	/* Throughput policy: v_rsq_f32 and FMA on the FP32 VALU, per-tile
	   partial sums combined with compensation, FP64 position drift,
	   deterministic tree-shaped FP64 reductions. */
	SF_PRECISION_FAST = 1
};

/* What the "acc" block holds. */
enum sf_acc_mode {
	/* acc_i = -G m_i sum_j m_j r_ij / (r^2+eps^2)^(3/2): the force, which the
	   reference's kick uses as an acceleration (src/gravity.c:808-810,
	   src/simulation.c:382).  Default. */
	SF_ACC_REFERENCE = 0,
	/* acc_i = -G sum_j m_j r_ij / (r^2+eps^2)^(3/2): the acceleration; the
	   SPH term likewise omits the m_i factor of src/sph.c:274-276. */
	SF_ACC_PHYSICAL = 1
};

/* Barnes-Hut flavour used by tree_build / solve_gravity_part_tree. */
enum sf_tree_mode {
	/* Reference octree: serial insertion in index order, leaf capacity 8,
	   open iff bodies > 8 and (mean width / distance to centre)^2 >= theta,
	   leaves used as monopoles with self-removal (src/tree.c:74-318,
	   src/gravity.c:273-660).  Bit-identical node array, acc and pot. */
	SF_TREE_REFERENCE = 0,
	/* GPU build from 63-bit Morton codes (FP64 quantisation), stable radix
	   sort, binary radix tree, fixed-order monopole aggregation; wavefront
	   coherent traversal with a group opening test and exact direct sums
	   over leaf particles. */
	SF_TREE_LBVH = 1
};

/* SPH formulation used by solve_sph. */
enum sf_sph_model {
	/* Gaussian kernel, polytrope P = (0.1 rho)^2, r^2 > 1 cutoff
	   (src/sph.c:14-283). */
	SF_SPH_REFERENCE = 0,
	/* Cubic spline (support radius h, W = 8/(pi h^3) {1-6q^2+6q^3; 2(1-q)^3}),
	   ideal gas P = (gamma-1) rho u with evolved specific internal energy u. */
	SF_SPH_CUBIC_IDEAL = 1
};

/* Pair search of the SF_SPH_CUBIC_IDEAL model. */
enum sf_sph_search {
	/* every pair is tested (the reference's particle-particle loops) */
	SF_SPH_SEARCH_ALL = 0,
	/* particles hashed into cubic cells of the largest support radius; each
	   particle tests the 27 cells around its own */
	SF_SPH_SEARCH_GRID = 1
};

/* Where sim.mem is authoritative between calls. */
enum sf_mem_mode {
	/* Every entry point uploads the blocks it reads and downloads the blocks
	   it writes, so host sim.mem is authoritative after each call returns,
	   as the reference's OpenMP offload maps do (src/gravity.c:688). */
	SF_MEM_MIRROR = 0,
	/* The device copy is authoritative; host sim.mem changes only through
	   sf_sync().  Entry points do no host transfers. */
	SF_MEM_RESIDENT = 1
};

typedef struct {
	double epsilon;              /* EPSILON, config.h:173 (1e-3) */
	double G;                    /* G, config.h:178 (1) */
	double dt;                   /* TIMESTEP_SIZE, config.h:170 (1) */
	double theta;                /* TREE_THETA, config.h:44 (0.5) */
	unsigned int tree_nodes_max; /* TREE_NODES_MAX, config.h:41 (262144) */
	unsigned int tree_depth_max; /* TREE_DEPTH_MAX, config.h:38 (16, at most 16) */
	unsigned int leaf_size;      /* SF_TREE_LBVH leaf capacity (3) */
	int enable_tree;             /* ENABLE_TREE, config.h:35 (1) */
	int enable_sph;              /* ENABLE_SPH, config.h:190 (0) */
	int kahan_solver;            /* ENABLE_KAHAN_SUMMATION_SOLVER, config.h:194 (1) */
	int kahan_energy;            /* ENABLE_KAHAN_SUMMATION_ENERGY, config.h:198 (1) */
	int precision;               /* enum sf_precision (EXACT) */
	int acc_mode;                /* enum sf_acc_mode (REFERENCE) */
	int tree_mode;               /* enum sf_tree_mode (REFERENCE) */
	int sph_model;               /* enum sf_sph_model (REFERENCE) */
	int mem_mode;                /* enum sf_mem_mode (MIRROR) */
	double sph_gamma;            /* SF_SPH_CUBIC_IDEAL adiabatic index (5/3) */
	int logging;                 /* write <out_dir>/statistics.csv and
	                                timings_sim_step.csv like the reference
	                                (config.h:120-123) (1) */
	int verbose;                 /* print energies per step (simulation.c:503) (0) */
	char out_dir[256];           /* OUTPUT_DIR, config.h:75 ("./out") */
	int quadrupole;              /* SF_TREE_LBVH cells carry traceless quadrupoles
	                                as well as monopoles (0: monopoles, as the
	                                reference's tree) */
	int sph_search;              /* enum sf_sph_search of the cubic model (GRID) */
	double mac_alpha;            /* SF_TREE_LBVH relative opening test (0: off, the
	                                default): a cell that passes the theta test is
	                                also required to satisfy G M l^2 <= mac_alpha
	                                |a|_min d^4 (quadrupole cells: G M l^3 <=
	                                mac_alpha |a|_min d^5, and those that also meet
	                                G M l^2 <= 0.2 mac_alpha |a|_min d^4 are
	                                summed as monopoles), with l its size, d its
	                                distance to the walk group's box and |a|_min the
	                                smallest previous acceleration in the group (sim_step
	                                reads it from the acc block; sim_solv and
	                                solve_gravity_part_tree first walk with the theta
	                                test alone).  Bounds the force error relative to
	                                each particle's acceleration. */
} sf_params_t;

/* Fill *p with the defaults listed above. */
void sf_params_default(sf_params_t* p);
/* Replace the process-wide parameters.  Fails if a live simulation was
   initialised with a different enable_sph (the sim.mem layout). */
int sf_set_params(const sf_params_t* p);
int sf_get_params(sf_params_t* p);

/* Message of the last failure in this thread ("" if none). */
const char* sf_last_error(void);

/* Library version string. */
const char* sf_version(void);

/* Number of reals per particle in sim.mem under the current parameters:
   13, or 16 with enable_sph (reference enum sim_conf). */
unsigned int sf_sim_dof(void);

enum sf_sync_dir { SF_SYNC_TO_DEVICE = 0, SF_SYNC_TO_HOST = 1 };
/* Copy the whole of sim.mem between host and device. */
int sf_sync(const sim_t* sim_ptr, int direction);

/* Device address of a sim.mem block (same layout as sim_find), for
   zero-copy interop.  NULL on error. */
void* sf_device_find(const sim_t* sim_ptr, enum sim_conf parameter_enum);

/* Extension per-particle arrays that are not part of sim.mem. */
enum sf_ext_array {
	SF_EXT_U = 0,      /* specific internal energy (SF_SPH_CUBIC_IDEAL), N floats */
	SF_EXT_DUDT = 1,   /* its time derivative, N floats */
	SF_EXT_X64 = 2     /* FP64 positions used by the FAST drift, 3N doubles */
};
void* sf_device_ext(const sim_t* sim_ptr, int which);
/* Host<->device copy of an extension array (count elements). */
int sf_ext_copy(const sim_t* sim_ptr, int which, void* host, int direction);

typedef struct {
	double T;        /* sum ken (FP64), as src/simulation.c:455-475 */
	double V;        /* sum pen (FP64), as src/simulation.c:477-497 */
	double E;        /* T + V, as logged by the reference */
	double W;        /* pair potential sum_{i<j} -G m_i m_j / s_ij = 0.5 sum_i m_i (pot_i + G m_i / eps) */
	double U;        /* sum m_i u_i (SF_SPH_CUBIC_IDEAL), else 0 */
	double P[3];     /* linear momentum sum m v (FP64) */
	double L[3];     /* angular momentum sum m x cross v (FP64) */
	double K;        /* 0.5 sum m |v|^2 of the current velocities (FP64); T is the
	                    sum of the per-particle energies a step last wrote */
} sf_stats_t;
/* Statistics of the current device state (the T, V of the last sim_step
   are the values it logged). */
int sf_get_stats(const sim_t* sim_ptr, sf_stats_t* out);

/* --- initial conditions (synthetic, seeded, FP64 on the host) ---
   Each generator writes host sim.mem (pos, vel, mas; rad is left to the
   caller) and, in SF_MEM_RESIDENT mode, then uploads it. */
/* Uniform ball of the given radius centred on the origin: r = radius *
   cbrt(u), isotropic direction (as src/initcond.c:73-86), velocity
   v = v_hubble * x, masses total_mass / N. */
int sf_ic_uniform_ball(sim_t* sim_ptr, uint32_t seed, double radius, double total_mass, double v_hubble);
/* Plummer sphere (scale a, mass M) on particles [first, first+count):
   radii by inverse CDF, velocities by Aarseth-Henon-Wielen rejection, then
   shifted by centre and bulk velocity.  Masses M / count. */
int sf_ic_plummer(sim_t* sim_ptr, uint32_t seed, unsigned int first, unsigned int count,
                  double a, double M, const double centre[3], const double bulk_v[3]);

/* Diagnostics of the SF_TREE_LBVH tree last built for tree_ptr (host
   outputs; any may be NULL): the sorted 63-bit Morton codes and the sorting
   permutation (N each); counts[4] = cell interactions and particle
   interactions of the last walk, the number of wavefronts that met a
   malformed tree (0 unless something is broken) and the number of 64-node
   batches the walk classified; node_cm = 2N-1 records
   (com x, y, z, mass), internal nodes first; node_lr_range = N-1 records
   (left, right, first, count), child ids >= N-1 being leaves of sorted
   particle id - (N-1); node_cell = 2N-1 records (centre x, y, z, side) of
   the box in each node's opening test: its particles' bounding box for
   nodes of at most 64 particles, its Morton cell above.  Any of them may be
   NULL. */
int sf_lbvh_export(const tree_t* tree_ptr, unsigned long long* keys, unsigned int* order,
                   unsigned long long counts[4], float* node_cm, int* node_lr_range, float* node_cell);

/* --- kernel timing with HIP events on the library stream --- */
void sf_prof_enable(int on);
void sf_prof_reset(void);
/* Total milliseconds and launch count of kernels recorded under `name`
   ("pp_fast", "pp_exact", "kdk", ...).  Synchronises the library stream. */
int sf_prof_query(const char* name, double* total_ms, unsigned long* count);

/* FP64 direct sum on the GPU for the listed particles (host index list, host
   outputs 3*n_idx and n_idx doubles), same acc_mode semantics as the
   solvers, self pair included.  A precision reference, not a solver. */
int sf_pp_direct_f64(const sim_t* sim_ptr, const unsigned int* idx, unsigned int n_idx,
                     double* acc_out, double* pot_out);
/* FP64 cubic-model SPH (SF_SPH_CUBIC_IDEAL) over every pair, for the listed
   particles of the current state (pos, vel, mas, rad, u): density, pressure
   acceleration (acc_mode semantics) and du/dt into host arrays of n_idx,
   3 n_idx and n_idx doubles.  A precision reference, not a solver. */
int sf_sph_direct_f64(const sim_t* sim_ptr, const unsigned int* idx, unsigned int n_idx,
                      double* rho_out, double* acc_out, double* dudt_out);

/* --- adjoint of the force and of the KDK step (device pointers) ---
   State: x, v are 3N floats (AoS), m is N floats, all device memory.
   The tape keeps (x, v, a) every `ckpt_every` steps on the device
   (48 N bytes each: x in FP64), and the backward recomputes ckpt_every - 1 steps per
   segment from them: ckpt_every = 1 trades 48 N bytes per step for no
   recomputation. */
typedef struct sf_tape sf_tape;
/* Run `nsteps` KDK steps of PP gravity from (x0, v0) with the current
   epsilon, G, dt and acc_mode; write the final state to (xN, vN) and return
   a tape for sf_adjoint_backward.  `stream` is the caller's hipStream_t
   (NULL: HIP's default null stream); the work is ordered on it and these
   calls return without waiting for it, except sf_force_dev and
   sf_force_vjp_dev, which synchronise it.  A tape's device memory is
   allocated (hipMalloc, which may synchronise) when it is made;
   sf_tape_free keeps the last freed tape's memory, and the next
   sf_adjoint_forward with the same N, nsteps and checkpoint interval (every
   iteration of a training loop after the first) reuses it without
   allocating, once that tape's queued work is done (a device-side wait). */
int sf_adjoint_forward(unsigned int N, const float* x0, const float* v0, const float* m,
                       unsigned int nsteps, unsigned int ckpt_every,
                       float* xN, float* vN, void* stream, sf_tape** tape_out);
/* Given dL/dxN and dL/dvN (device, 3N each; either may be NULL for zero),
   write dL/dx0 and dL/dv0 (device, 3N each). */
int sf_adjoint_backward(sf_tape* tape, const float* gxN, const float* gvN,
                        float* gx0, float* gv0, void* stream);
void sf_tape_free(sf_tape* tape);
/* One force evaluation on device arrays: acc (3N) from x (3N) and m (N). */
int sf_force_dev(unsigned int N, const float* x, const float* m, float* acc, float* pot, void* stream);
/* Force vector-Jacobian product: g = (d acc / d x)^T mu, device arrays. */
int sf_force_vjp_dev(unsigned int N, const float* x, const float* m, const float* mu, float* g, void* stream);
/* sf_force_dev without allocation.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under stream
   capture, allocation-free serving loops).  Same kernels and summation
   order as sf_force_dev, so the same acc and pot bit for bit; its scratch
   (packed bodies and j-slice partials, 8.3 MiB) is a static device array
   that each GPU the library runs on holds until the process ends.  N above
   SF_NOALLOC_MAX_N fails with an error.  The work is queued on `stream`
   and the call returns without waiting for it.  Every call on one GPU
   shares that GPU's scratch: calls that overlap on one GPU (other streams,
   other host threads) race. */
#define SF_NOALLOC_MAX_N 16384u
int sf_force_dev_noalloc(unsigned int N, const float* x, const float* m, float* acc, float* pot, void* stream);

/* --- the reference program's run loop (src/main.c:348-611) ---
   The reference sets these with config.h macros; the physics parameters
   come from sf_set_params, and every path is under its out_dir
   (OUTPUT_DIR). */
typedef struct {
	unsigned int num_bodies;        /* NUM_BODIES, config.h:161 */
	unsigned int num_timesteps;     /* NUM_TIMESTEPS, config.h:22: steps 0 .. num_timesteps */
	unsigned int output_interval;   /* OUTPUT_INTERVAL, config.h:78; 0 = undefined (every step) */
	int enable_sim;                 /* ENABLE_SIM, config.h:26 */
	int enable_vis;                 /* ENABLE_VIS, config.h:29 */
	const char* sim_filename;       /* SIM_FILENAME, config.h:85 ("sim/step_%04u"; ".raw" is
	                                   appended); NULL = undefined: no snapshots */
	const char* vis_filename;       /* VIS_FILENAME, config.h:88 ("vis/step_%04u"; the file
	                                   type's extension is appended); NULL = no frames */
	image_filetype_t vis_file_type; /* VIS_FILE_FORMAT, config.h:108 */
	unsigned int vis_sizex, vis_sizey; /* visualization_dimensions, src/main.c:81-99 */
	const char* init_cond_file;     /* INIT_COND_FILE, config.h:82; NULL = initcond_generate */
	int verbose;                    /* print the per-step progress lines of src/main.c:424-425 */
} sf_run_config_t;

/* Replaces the body of the reference's main() (src/main.c:348-611): set up
   the visualisation and the simulation, apply the initial conditions (read
   init_cond_file; else initcond_generate and one sim_solv), then for each
   step: stop if <out_dir>/stop exists (and remove it), set step_number,
   save the snapshot (before the step; or read it when enable_sim is 0),
   vis_draw, sim_step (all but the last step), un-premultiply and vis_save.
   With enable_sim 0 and an output interval, steps advance by the interval.
   A file name pattern holds exactly one %u conversion, which receives
   step / output_interval (or the step).  Returns 0 when the run ends or
   stops, 1 when it ends early because a replayed snapshot could not be read
   or a file name did not fit (the reference's program then still exits with
   success), -1 when the set-up fails. */
int sf_run(const sf_run_config_t* cfg);

#ifdef __cplusplus
}
#endif

#endif /* STARFLOOD_EXT_H */
