/*
 * sf_oracle.h -- CPU oracle for the force-and-integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, as the checker.
 *
 * Two kinds of functions:
 *  - orc_*_ref: restatements of the reference's FP32 algorithms in its
 *    operation order (each cites the reference file:line it follows), which
 *    the tests pin against golden fixtures produced by the reference itself;
 *  - orc_*_f64: FP64 statements of the mathematics (direct sums, the
 *    cubic-spline SPH model, the adjoint), pinned by closed forms,
 *    invariants and finite differences.
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>

/* pcg4d (reference src/rng.h:13-40) */
void orc_pcg4d(uint32_t s[4]);

/* PP gravity (reference src/gravity.c:667-819), FP32, per-pair Kahan.
   acc_phys: 0 = acc_i = -G m_i sum (reference), 1 = -G sum. */
void orc_pp_ref(unsigned N, const float* pos, const float* mas, double eps, double G, int kahan_acc, int kahan_pot,
                int acc_phys, float* acc, float* pot);

/* SPH, Gaussian kernel + polytrope (reference src/sph.c:14-283):
   writes rho, prs and adds the pressure force into acc. */
void orc_sph_ref(unsigned N, const float* pos, const float* mas, const float* rad, double eps, int kahan,
                 float* rho, float* prs, float* acc);

/* Reference octree node (src/tree.h:16-25). */
typedef struct {
	uint32_t parent, bodies, branch[8];
	float bounds[6];
	float params[4];
} orc_node;

/* Tree bounds of src/simulation.c:245-301 for a step number. */
void orc_tree_bounds_ref(unsigned N, const float* pos, unsigned step_number, float bounds[6]);
/* Serial build (src/tree.c:74-318).  Returns 0, or -1 (out of nodes) / -2
   (too deep) with the partial tree, as the reference leaves it. */
int orc_tree_build_ref(unsigned N, const float* pos, const float* mas, const float bounds[6], unsigned max_nodes,
                       int depth_max, orc_node* node, unsigned* num_nodes);
/* Walk (src/gravity.c:245-663). */
void orc_tree_walk_ref(unsigned N, const float* pos, const float* mas, const orc_node* node, unsigned node_cap,
                       unsigned iter_max, double eps, double G, double theta, int kahan_acc, int kahan_pot,
                       int acc_phys, float* acc, float* pot);

/* A simulation as the reference holds it: sim.mem blocks of N floats. */
typedef struct {
	unsigned N, dof, step_number;
	float* mem;
	double eps, G, dt, theta;
	int enable_tree, enable_sph, kahan_solver, kahan_energy, acc_phys;
	unsigned tree_nodes_max;
	int tree_depth_max;
	orc_node* node;          /* tree_nodes_max nodes when enable_tree */
	unsigned num_nodes;
	int tree_status;
	double T, V;             /* of the last orc_sim_step */
} orc_sim;

/* sim_solv (src/simulation.c:231-345) */
int orc_sim_solv(orc_sim* s);
/* sim_step (src/simulation.c:349-519), FP32 KDK, FP64 Kahan T and V */
int orc_sim_step(orc_sim* s);

/* FP64 direct sum for listed particles (self pair included) */
void orc_pp_f64(unsigned N, const float* pos, const float* mas, double eps, double G, int acc_phys,
                const unsigned* idx, unsigned n_idx, double* acc, double* pot);

/* Cubic-spline SPH with an ideal gas, FP64 (no reference counterpart):
   rho_i = sum_j m_j Wbar_ij, P = (gamma-1) rho u,
   a_i   = -sum_j m_j (P_i/rho_i^2 + P_j/rho_j^2) grad_i Wbar_ij,
   du_i  = 1/2 sum_j m_j (P_i/rho_i^2 + P_j/rho_j^2) v_ij . grad_i Wbar_ij,
   Wbar_ij = (W(r, h_i) + W(r, h_j)) / 2, W the M4 spline of support h. */
void orc_sph_cubic_f64(unsigned N, const double* pos, const double* vel, const double* mas, const double* h,
                       const double* u, double gamma, double* rho, double* prs, double* acc, double* dudt);
/* the same for the listed rows only: rho and P of all N (the force rows
   need their neighbours'), acceleration (3 n_rows) and du/dt (n_rows) */
void orc_sph_cubic_rows_f64(unsigned N, const double* pos, const double* vel, const double* mas, const double* h,
                            const double* u, double gamma, unsigned n_rows, const unsigned* rows, double* rho,
                            double* prs, double* acc_rows, double* dudt_rows);
double orc_cubic_w(double r, double h);
double orc_cubic_dw(double r, double h);

/* FP64 PP acceleration (acc_phys semantics as above) of double positions. */
void orc_force_f64(unsigned N, const double* x, const double* m, double eps, double G, int acc_phys, double* acc);
/* g = (d acc / d x)^T mu, FP64 */
void orc_force_vjp_f64(unsigned N, const double* x, const double* m, double eps, double G, int acc_phys,
                       const double* mu, double* g);
/* nsteps KDK steps in FP64 from (x, v), in place. */
void orc_kdk_f64(unsigned N, double* x, double* v, const double* m, double eps, double G, double dt, int acc_phys,
                 unsigned nsteps);
/* Gradient of L = mean((x_N - target)^2) with respect to x0 and v0 by the
   discrete adjoint of the FP64 KDK; returns L. */
double orc_kdk_adjoint_f64(unsigned N, const double* x0, const double* v0, const double* m, double eps, double G,
                           double dt, int acc_phys, unsigned nsteps, const double* target, double* gx0, double* gv0);

/* 63-bit Morton codes of the SF_TREE_LBVH build, by a bit-by-bit
   interleave (no reference counterpart): FP64 bounding box, extent = longest
   side * (1 + 1e-9), q = floor(clamp((x - lo) * (2^21 / extent), 0,
   2^21 - 1)), bit b of q_x at 3b + 2, of q_y at 3b + 1, of q_z at 3b. */
void orc_morton63(unsigned N, const float* pos, unsigned long long* key);

/* wall-clock seconds (CLOCK_MONOTONIC, as the reference's timing.h) */
double orc_now(void);
int orc_num_threads(void);

#endif
