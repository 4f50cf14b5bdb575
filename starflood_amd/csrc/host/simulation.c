/*
 * simulation.c -- sim_find / sim_init / sim_free / sim_solv / sim_step with
 * the reference's signatures and semantics (src/simulation.c), driving the
 * device kernels on the simulation's stream.
 */
#define _POSIX_C_SOURCE 200112L

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sf_host.h"

#define SF_ALIGNMENT 4194304u   /* STARFLOOD_ALIGNMENT, config.h:230 */
/* smaller sim.mem stays ordinary memory: copies from page-locked memory go
   through the DMA engine, whose per-copy latency costs more than its rate
   gains on a few-MB state (N = 65536: reference-tree solve 3.2 -> 4.0 ms),
   while a 13.6 MB state (N = 262144) steps 2.7 % faster in mirror mode */
#define SF_PIN_MIN_BYTES 8388608u

/* statistics.csv and timings_sim_step.csv (src/simulation.c:24-30,58-76) */
static FILE* g_log_statistics = NULL;
static FILE* g_log_timings = NULL;
static hipEvent_t g_tev[7];
static int g_tev_ready = 0;

/* reference src/simulation.c:33-53 */
real* sim_find(const sim_t* SF_RESTRICT sim_ptr, enum sim_conf parameter_enum) {
	const unsigned int dof = sf_sim_dof();
	unsigned int e = (unsigned int)parameter_enum;
	if (!sim_ptr || e >= dof) return (real*)NULL;
	if (e < 3u) e = SF_E_POS;
	else if (e < 6u) e = SF_E_VEL;
	else if (e < 9u) e = SF_E_ACC;
	return &((real*)sim_ptr->mem)[(size_t)sim_ptr->N * (size_t)e];
}

static FILE* open_log(const char* name, const char* header) {
	char path[512];
	snprintf(path, sizeof(path), "%s/%s", sf_params()->out_dir, name);
	FILE* f = fopen(path, "w");
	if (!f) {
		fprintf(stderr, "%s error: fopen(\"%s\", \"%s\") ", "log_init()", path, "w");
		perror("failed");
		return NULL;
	}
	fputs(header, f);
	fflush(f);
	return f;
}

static void close_logs(void) {
	if (g_log_statistics) fclose(g_log_statistics);
	if (g_log_timings) fclose(g_log_timings);
	g_log_statistics = g_log_timings = NULL;
}

/* reference src/simulation.c:55-187 */
int sim_init(sim_t* SF_RESTRICT sim_ptr, unsigned int N) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	if (p->logging) {
		close_logs();
		g_log_statistics = open_log("statistics.csv", "step_number,total_energy,potential_energy,kinetic_energy\n");
		if (!g_log_statistics) {
			fprintf(stderr, "%s error: %s failed!\n", "sim_init()", "log_init(&log_statistics)");
			return STARFLOOD_FAILURE;
		}
		g_log_timings = open_log("timings_sim_step.csv",
		                         "step_number,kick_0,drift,sim_solv,kick_1,ken_sum,pen_sum\n");
		if (!g_log_timings) {
			fprintf(stderr, "%s error: %s failed!\n", "sim_init()", "log_init(&log_timings_sim_step)");
			close_logs();
			return STARFLOOD_FAILURE;
		}
	}

	sim_t sim = *sim_ptr;
	const unsigned int dof = sf_sim_dof();
	const size_t mem_size = sizeof(real) * (size_t)N * (size_t)dof;
	/* page-locked (the mirror-mode copies then run at full PCIe rate), still
	   SF_ALIGNMENT aligned inside the allocation; ordinary aligned memory,
	   as the reference's, for small systems or when pinning fails */
	void* mem = NULL;
	void* pinned = NULL;
	if (mem_size >= SF_PIN_MIN_BYTES &&
	    hipHostMalloc(&pinned, mem_size + SF_ALIGNMENT, hipHostMallocDefault) == hipSuccess && pinned) {
		mem = (void*)(((uintptr_t)pinned + (SF_ALIGNMENT - 1u)) & ~(uintptr_t)(SF_ALIGNMENT - 1u));
	} else {
		(void)hipGetLastError();
		pinned = NULL;
		if (posix_memalign(&mem, (size_t)SF_ALIGNMENT, mem_size ? mem_size : 1u) != 0 || !mem) {
			fprintf(stderr, "%s error: mem is NULL after posix_memalign(&mem, %zu, %zu)\n", "sim_init()",
			        (size_t)SF_ALIGNMENT, mem_size);
			close_logs();
			return STARFLOOD_FAILURE;
		}
	}
	memset(mem, 0, mem_size);

	sf_ctx* c = sf_ctx_create(mem, N, dof);
	if (!c) {
		if (pinned) hipHostFree(pinned);
		else free(mem);
		close_logs();
		return STARFLOOD_FAILURE;
	}
	c->host_pinned_base = pinned;

	if (p->verbose) {
		printf("Simulation Memory Addresses:\n  mem: %p (%zu bytes)\n\n", mem, mem_size);
	}

	/* src/simulation.c:165-174: a tree_init failure is reported, not fatal */
	if (STARFLOOD_SUCCESS != tree_init(&sim.tree, p->tree_nodes_max)) {
		fprintf(stderr, "%s error: tree_init() failed.\n", "sim_init()");
	}

	sim.step_number = 0u;
	sim.N = N;
	sim.mem = mem;
	*sim_ptr = sim;
	return STARFLOOD_SUCCESS;
}

/* reference src/simulation.c:189-228.  Unlike the reference (which leaks
   it), the tree is released too and sim.tree.node cleared. */
int sim_free(sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	close_logs();
	sim_t sim = *sim_ptr;
	void* pinned = NULL;
	if (sim.mem) {
		sf_ctx* c = sf_ctx_lookup(sim.mem, "sim_free");
		if (c) {
			pinned = c->host_pinned_base;
			sf_ctx_destroy(c);   /* waits for the context's stream */
		}
	}
	if (pinned) hipHostFree(pinned);
	else free(sim.mem);
	if (sim.tree.node) tree_free(&sim.tree);
	sim.step_number = 0u;
	sim.N = 0u;
	sim.mem = NULL;
	*sim_ptr = sim;
	return STARFLOOD_SUCCESS;
}

/* blocks a solve reads / writes in mirror mode */
static int upload_solver_inputs(sf_ctx* c) {
	if (sf_upload_blocks(c, SF_E_POS, 3u)) return -1;
	if (sf_upload_blocks(c, SF_E_MAS, 1u)) return -1;
	if (sf_params()->enable_sph && sf_upload_blocks(c, SF_E_RAD, 1u)) return -1;
	return 0;
}

static int download_solver_outputs(sf_ctx* c) {
	if (sf_download_blocks(c, SF_E_ACC, 3u)) return -1;
	if (sf_download_blocks(c, SF_E_POT, 1u)) return -1;
	if (sf_params()->enable_sph && sf_download_blocks(c, SF_E_RHO, 2u)) return -1;
	return 0;
}

/* Solver sequence of src/simulation.c:231-345 on device state. */
int sf_solve_all(sf_ctx* c, tree_t* tree, unsigned int step_number, float sph_pred_dt, int prev_acc) {
	const sf_params_t* p = sf_params();
	int rc = 0;
	if (c->N == 0) return 0;   /* no particles: no forces, no densities */
	if (p->enable_sph && sf_sph_prepare(c, sph_pred_dt)) return -1;
	if (p->enable_tree) {
		if (sf_solve_gravity_tree(c, tree, step_number, prev_acc)) rc = -1;
	} else {
		if (sf_solve_gravity_pp(c)) rc = -1;
	}
	if (p->enable_sph) {
		if (sf_solve_sph(c)) rc = -1;
	}
	return rc;
}

/* reference src/simulation.c:231-345.  A failed tree build fails the call
   (the reference ignores tree_build's return code, :312). */
int sf_sim_solv(sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sim_solv");
	if (!c) return STARFLOOD_FAILURE;
	int rc = 0;
	if (upload_solver_inputs(c)) return STARFLOOD_FAILURE;
	if (sf_solve_all(c, &sim_ptr->tree, sim_ptr->step_number, 0.0f, 0)) rc = -1;
	if (download_solver_outputs(c)) rc = -1;
	if (sf_finish(c, "sim_solv")) rc = -1;
	if (sf_params()->enable_tree && c->N && sf_tree_check(&sim_ptr->tree, c->stream, "sim_solv")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

static void tev_record(int k, hipStream_t s) {
	if (g_tev_ready) hipEventRecord(g_tev[k], s);
}

/* reference src/simulation.c:349-519 */
int sf_sim_step(sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sim_step");
	if (!c) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	sim_t sim = *sim_ptr;
	const unsigned int N = c->N;
	const unsigned int n3 = 3u * N;
	hipStream_t s = c->stream;
	float* pos = sf_blk(c->d_mem, N, SF_E_POS);
	float* vel = sf_blk(c->d_mem, N, SF_E_VEL);
	float* acc = sf_blk(c->d_mem, N, SF_E_ACC);
	float* mas = sf_blk(c->d_mem, N, SF_E_MAS);
	float* pot = sf_blk(c->d_mem, N, SF_E_POT);
	float* ken = sf_blk(c->d_mem, N, SF_E_KEN);
	float* pen = sf_blk(c->d_mem, N, SF_E_PEN);
	const float dt = (float)p->dt;          /* const real dt = (real)TIMESTEP_SIZE */
	const float half_dt = 0.5f * dt;       /* (real)0.5 * dt */
	const int exact = p->precision == SF_PRECISION_EXACT;
	const int timing = p->logging && g_log_timings;
	/* the cubic SPH model carries u: it is kicked with v, and the solve
	   predicts v and u to the end of the step from the previous forces */
	const int cubic = p->enable_sph && p->sph_model == SF_SPH_CUBIC_IDEAL;
	int rc = 0;

	if (timing && !g_tev_ready) {
		g_tev_ready = 1;
		for (int k = 0; k < 7; ++k)
			if (hipEventCreate(&g_tev[k]) != hipSuccess) g_tev_ready = 0;
	}

	/* mirror mode: pos, vel, acc, mas (+ rad) are read */
	if (sf_upload_blocks(c, SF_E_POS, 10u)) return STARFLOOD_FAILURE;
	if (p->enable_sph && sf_upload_blocks(c, SF_E_RAD, 1u)) return STARFLOOD_FAILURE;

	if (cubic && (sf_ctx_ensure_u(c) || sfd_sph_u_kick(c->d_u, c->d_dudt, N, half_dt, s))) return STARFLOOD_FAILURE;
	if (timing) tev_record(0, s);
	if (exact) {
		if (sfd_kick_exact(vel, acc, n3, half_dt, s)) return STARFLOOD_FAILURE;    /* kick_0 */
		if (timing) tev_record(1, s);
		if (sfd_drift_exact(pos, vel, n3, dt, s)) return STARFLOOD_FAILURE;       /* drift */
	} else {
		/* x64 restarts from pos where a host-side change left them apart */
		if (sfd_kick_drift_fast(vel, acc, c->d_x64, pos, n3, half_dt, p->dt, 1, s)) return STARFLOOD_FAILURE;
		if (timing) tev_record(1, s);
	}
	if (timing) tev_record(2, s);

	/* acc still holds the force that kick_0 used: the previous evaluation */
	if (sf_solve_all(c, &sim_ptr->tree, sim.step_number, cubic ? half_dt : 0.0f, 1)) {
		fprintf(stderr, "%s error: %s failed.\n", "sim_step()", "sim_solv()");
		rc = -1;
	}
	if (timing) tev_record(3, s);

	if (exact) {
		if (sfd_kick_exact(vel, acc, n3, half_dt, s)) return STARFLOOD_FAILURE;    /* kick_1 */
	} else {
		/* kick_1 and the energy terms below, in one pass */
		if (sfd_kick_energy_fast(vel, acc, mas, pot, N, half_dt, ken, pen, s)) return STARFLOOD_FAILURE;
	}
	if (cubic && sfd_sph_u_kick(c->d_u, c->d_dudt, N, half_dt, s)) return STARFLOOD_FAILURE;
	if (timing) tev_record(4, s);
	if (exact && sfd_energy_terms(vel, mas, pot, N, ken, pen, s)) return STARFLOOD_FAILURE;

	double* red_out = c->d_red + (size_t)(sfd_reduce_scratch_doubles(N) - SF_RED_VALUES);
	if (exact) {
		if (sfd_energy_sum_sequential(ken, N, red_out + 12, s)) return STARFLOOD_FAILURE;   /* ken_sum */
		if (timing) tev_record(5, s);
		if (sfd_energy_sum_sequential(pen, N, red_out + 13, s)) return STARFLOOD_FAILURE;   /* pen_sum */
		if (timing) tev_record(6, s);
	}
	/* FP64 W, P, L (and T, V in the fast policy) */
	if (sfd_reduce_state(pos, exact ? NULL : c->d_x64, vel, mas, pot, ken, pen, c->have_u ? c->d_u : NULL, N, p->G,
	                     p->epsilon, c->d_red, s))
		return STARFLOOD_FAILURE;
	if (!exact && timing) {
		tev_record(5, s);
		tev_record(6, s);
	}

	/* mirror mode: pos, vel, acc, mas, pot, ken, pen (+ rho, prs) are written */
	if (sf_download_blocks(c, SF_E_POS, 13u)) return STARFLOOD_FAILURE;
	if (p->enable_sph && sf_download_blocks(c, SF_E_RHO, 2u)) return STARFLOOD_FAILURE;

	const int need_host = p->mem_mode == SF_MEM_MIRROR || p->logging || p->verbose;
	if (need_host) {
		SF_HIP(hipMemcpyAsync(c->h_red, red_out, SF_RED_VALUES * sizeof(double), hipMemcpyDeviceToHost, s),
		       "sim_step");
		if (sf_finish(c, "sim_step")) return STARFLOOD_FAILURE;
		/* the step synchronises here anyway: report a tree walk cut short */
		if (p->enable_tree && N && sf_tree_check(&sim_ptr->tree, s, "sim_step")) rc = -1;
		sf_stats_t* st = &c->stats;
		const double* r = c->h_red;
		st->T = exact ? r[12] : r[0];
		st->V = exact ? r[13] : r[1];
		st->E = st->V + st->T;
		st->W = r[2];
		st->U = r[3];
		for (int k = 0; k < 3; ++k) {
			st->P[k] = r[4 + k];
			st->L[k] = r[7 + k];
		}
		st->K = r[10];
		if (p->logging && g_log_statistics) {
			fprintf(g_log_statistics, "%u,%.015f,%.015f,%.015f\n", sim.step_number, st->V + st->T, st->V, st->T);
			fflush(g_log_statistics);
		}
		if (timing && g_tev_ready) {
			float ms[6] = {0, 0, 0, 0, 0, 0};
			for (int k = 0; k < 6; ++k) hipEventElapsedTime(&ms[k], g_tev[k], g_tev[k + 1]);
			fprintf(g_log_timings, "%u", sim.step_number);
			for (int k = 0; k < 6; ++k) fprintf(g_log_timings, ",%.03f", 1.0e3 * (double)ms[k]);
			fprintf(g_log_timings, "\n");
			fflush(g_log_timings);
		}
		if (p->verbose) {
			printf("E (tot) = % .015f\nV (pot) = % .015f\nT (kin) = % .015f\n", st->V + st->T, st->V, st->T);
		}
	}

	/* the reference writes its local copy back (src/simulation.c:516) */
	sim.tree = sim_ptr->tree;
	*sim_ptr = sim;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

int sf_get_stats(const sim_t* sim_ptr, sf_stats_t* out) {
	if (!sim_ptr || !out) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_get_stats");
	if (!c) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	const int exact = p->precision == SF_PRECISION_EXACT;
	hipStream_t s = c->stream;
	if (p->mem_mode == SF_MEM_MIRROR && sf_upload_blocks(c, SF_E_POS, 13u)) return STARFLOOD_FAILURE;
	float* ken = sf_blk(c->d_mem, N, SF_E_KEN);
	float* pen = sf_blk(c->d_mem, N, SF_E_PEN);
	double* red_out = c->d_red + (size_t)(sfd_reduce_scratch_doubles(N) - SF_RED_VALUES);
	if (!exact && sfd_x64_refresh(sf_blk(c->d_mem, N, SF_E_POS), c->d_x64, 3u * N, s)) return STARFLOOD_FAILURE;
	if (sfd_reduce_state(sf_blk(c->d_mem, N, SF_E_POS), exact ? NULL : c->d_x64, sf_blk(c->d_mem, N, SF_E_VEL),
	                     sf_blk(c->d_mem, N, SF_E_MAS), sf_blk(c->d_mem, N, SF_E_POT), ken, pen,
	                     c->have_u ? c->d_u : NULL, N, p->G, p->epsilon, c->d_red, s))
		return STARFLOOD_FAILURE;
	if (exact) {
		if (sfd_energy_sum_sequential(ken, N, red_out + 12, s)) return STARFLOOD_FAILURE;
		if (sfd_energy_sum_sequential(pen, N, red_out + 13, s)) return STARFLOOD_FAILURE;
	}
	SF_HIP(hipMemcpyAsync(c->h_red, red_out, SF_RED_VALUES * sizeof(double), hipMemcpyDeviceToHost, s),
	       "sf_get_stats");
	if (sf_finish(c, "sf_get_stats")) return STARFLOOD_FAILURE;
	const double* r = c->h_red;
	out->T = exact ? r[12] : r[0];
	out->V = exact ? r[13] : r[1];
	out->E = out->T + out->V;
	out->W = r[2];
	out->U = r[3];
	for (int k = 0; k < 3; ++k) {
		out->P[k] = r[4 + k];
		out->L[k] = r[7 + k];
	}
	out->K = r[10];
	/* a tree walk cut short since the last check fails the call */
	if (p->enable_tree && N && sf_tree_check((tree_t*)&sim_ptr->tree, s, "sf_get_stats")) return STARFLOOD_FAILURE;
	return STARFLOOD_SUCCESS;
}

/* The exported entry points, in the prototype form the caller's SOLVER_VIS
   selects (reference src/simulation.h:90-101).  The SOLVER_VIS form's render
   buffer is the reference's tree debug overlay, which is not drawn. */
#ifdef SOLVER_VIS
int sim_solv(sim_t* SF_RESTRICT sim_ptr, f32* SF_RESTRICT render_buffer, unsigned int sizex, unsigned int sizey) {
	(void)render_buffer;
	(void)sizex;
	(void)sizey;
	return sf_sim_solv(sim_ptr);
}

int sim_step(sim_t* SF_RESTRICT sim_ptr, f32* SF_RESTRICT render_buffer, unsigned int sizex, unsigned int sizey) {
	(void)render_buffer;
	(void)sizex;
	(void)sizey;
	return sf_sim_step(sim_ptr);
}
#else
int sim_solv(sim_t* SF_RESTRICT sim_ptr) { return sf_sim_solv(sim_ptr); }

int sim_step(sim_t* SF_RESTRICT sim_ptr) { return sf_sim_step(sim_ptr); }
#endif
