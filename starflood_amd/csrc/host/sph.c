/*
 * sph.c -- solve_sph with the reference's signature (src/sph.h:6,
 * src/sph.c:14-283): density, equation of state, pressure force added into
 * acc.  The model is a parameter (sf_params_t.sph_model): the reference's
 * Gaussian kernel and polytrope, or a cubic spline with an ideal gas whose
 * specific internal energy u (sf_device_ext SF_EXT_U) is integrated by
 * sim_step.
 */
#include <stdio.h>

#include "sf_host.h"

/* kernels in device/sph.hip */
size_t sfd_sph_tmp_floats(unsigned int n_pad);
int sfd_sph_ref(const float* pos, const float* mas, const float* rad, unsigned int N, int kahan, float* rho, float* prs,
                float* acc, float* tmp, hipStream_t s);
int sfd_sph_cubic_pack(const float* pos, const float* vel, const float* acc, const float* mas, const float* rad,
                       const float* u, const float* dudt, unsigned int N, unsigned int n_pad, float pred_dt, float* tmp,
                       hipStream_t s);
int sfd_sph_cubic(unsigned int N, unsigned int n_pad, float gamma, int acc_phys, float* rho, float* prs, float* acc,
                  float* dudt, float* tmp, float* part, unsigned int slices, hipStream_t s);
/* device/sph_grid.hip */
size_t sfd_sph_grid_bytes(unsigned int N);
int sfd_sph_cubic_grid(const float* rad, unsigned int N, unsigned int n_pad, float gamma, int acc_phys, float* rho,
                       float* prs, float* acc, float* dudt, float* tmp, void* scratch, hipStream_t s);

static int ensure_tmp(sf_ctx* c) {
	if (c->d_sph_tmp) return 0;
	SF_HIP(hipMalloc((void**)&c->d_sph_tmp, sfd_sph_tmp_floats(c->n_pad) * sizeof(float)), "solve_sph");
	return 0;
}

int sf_sph_prepare(sf_ctx* c, float pred_dt) {
	const sf_params_t* p = sf_params();
	if (!p->enable_sph || p->sph_model != SF_SPH_CUBIC_IDEAL || c->N == 0) return 0;
	if (ensure_tmp(c) || sf_ctx_ensure_u(c)) return -1;
	const unsigned int N = c->N;
	return sfd_sph_cubic_pack(sf_blk(c->d_mem, N, SF_E_POS), sf_blk(c->d_mem, N, SF_E_VEL),
	                          sf_blk(c->d_mem, N, SF_E_ACC), sf_blk(c->d_mem, N, SF_E_MAS),
	                          sf_blk(c->d_mem, N, SF_E_RAD), c->d_u, c->d_dudt, N, c->n_pad, pred_dt, c->d_sph_tmp,
	                          c->stream);
}

int sf_solve_sph(sf_ctx* c) {
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	if (N == 0) return 0;
	if (ensure_tmp(c)) return -1;
	float* rho = sf_blk(c->d_mem, N, SF_E_RHO);
	float* prs = sf_blk(c->d_mem, N, SF_E_PRS);
	float* acc = sf_blk(c->d_mem, N, SF_E_ACC);
	if (p->sph_model == SF_SPH_CUBIC_IDEAL && p->sph_search == SF_SPH_SEARCH_GRID) {
		if (sf_ctx_ensure_u(c)) return -1;
		if (!c->d_sph_grid) SF_HIP(hipMalloc(&c->d_sph_grid, sfd_sph_grid_bytes(N)), "solve_sph");
		return sfd_sph_cubic_grid(sf_blk(c->d_mem, N, SF_E_RAD), N, c->n_pad, (float)p->sph_gamma,
		                          p->acc_mode == SF_ACC_PHYSICAL, rho, prs, acc, c->d_dudt, c->d_sph_tmp,
		                          c->d_sph_grid, c->stream);
	}
	if (p->sph_model == SF_SPH_CUBIC_IDEAL) {
		/* slice partials reuse the PP kernel's scratch (same stream, after it) */
		const unsigned int slices = sf_pp_slice_count(c->n_pad);
		if (sf_ctx_ensure_u(c) || sf_ctx_reserve_part(c, slices)) return -1;
		return sfd_sph_cubic(N, c->n_pad, (float)p->sph_gamma, p->acc_mode == SF_ACC_PHYSICAL, rho, prs, acc,
		                     c->d_dudt, c->d_sph_tmp, c->d_part, slices, c->stream);
	}
	const int tok = sf_prof_begin("sph_ref", c->stream);
	const int rc = sfd_sph_ref(sf_blk(c->d_mem, N, SF_E_POS), sf_blk(c->d_mem, N, SF_E_MAS),
	                           sf_blk(c->d_mem, N, SF_E_RAD), N, p->kahan_solver, rho, prs, acc, c->d_sph_tmp,
	                           c->stream);
	sf_prof_end(tok, c->stream);
	return rc;
}

/* src/sph.c:14-283: reads pos, mas, rad (and, for the cubic model, vel and
   u as they are), writes rho and prs, adds the pressure force into acc. */
int solve_sph(const sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "solve_sph");
	if (!c) return STARFLOOD_FAILURE;
	if (!sf_params()->enable_sph) {
		sf_error("solve_sph", "enable_sph is 0 (sim.mem has no rad/rho/prs blocks)");
		return STARFLOOD_FAILURE;
	}
	/* pos, vel, acc, mas, then rad */
	if (sf_upload_blocks(c, SF_E_POS, 10u) || sf_upload_blocks(c, SF_E_RAD, 1u)) return STARFLOOD_FAILURE;
	int rc = sf_sph_prepare(c, 0.0f);
	if (!rc) rc = sf_solve_sph(c);
	if (sf_download_blocks(c, SF_E_ACC, 3u) || sf_download_blocks(c, SF_E_RHO, 2u)) rc = -1;
	if (sf_finish(c, "solve_sph")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
