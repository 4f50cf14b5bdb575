/*
 * sph.c -- solve_sph with the reference's signature (src/sph.h:6,
 * src/sph.c:14-283): density, equation of state, pressure force added into
 * acc.  The model is a parameter (sf_params_t.sph_model): the reference's
 * Gaussian kernel and polytrope, or a cubic spline with an ideal gas whose
 * specific internal energy u (sf_device_ext SF_EXT_U) is integrated by
 * sim_step.
 */
#include <stdio.h>
#include <stdlib.h>

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

/* device/sph_f64.hip */
int sfd_sph_cubic_f64(const float* pos, const float* vel, const float* mas, const float* rad, const float* u,
                      unsigned int N, double gamma, const unsigned int* idx, unsigned int n_idx, int acc_phys,
                      double* rho, double* cpr, double* out, hipStream_t s);

/* FP64 cubic-model SPH for listed particles of the current state: the
   precision reference of the FP32 passes (no reference counterpart). */
int sf_sph_direct_f64(const sim_t* sim_ptr, const unsigned int* idx, unsigned int n_idx, double* rho_out,
                      double* acc_out, double* dudt_out) {
	static const char* fn = "sf_sph_direct_f64";
	if (!sim_ptr || (!idx && n_idx) || (n_idx && (!rho_out || !acc_out || !dudt_out))) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	if (!p->enable_sph || p->sph_model != SF_SPH_CUBIC_IDEAL) {
		sf_error(fn, "needs enable_sph with sph_model SF_SPH_CUBIC_IDEAL");
		return STARFLOOD_FAILURE;
	}
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, fn);
	if (!c) return STARFLOOD_FAILURE;
	const unsigned int N = c->N;
	if (n_idx == 0 || N == 0) return STARFLOOD_SUCCESS;
	for (unsigned int k = 0; k < n_idx; ++k) {
		if (idx[k] >= N) {
			sf_error(fn, "index %u out of range (N = %u)", idx[k], N);
			return STARFLOOD_FAILURE;
		}
	}
	if (sf_upload_blocks(c, SF_E_POS, 2u * 3u) || sf_upload_blocks(c, SF_E_MAS, 1u) ||
	    sf_upload_blocks(c, SF_E_RAD, 1u) || sf_ctx_ensure_u(c))
		return STARFLOOD_FAILURE;
	unsigned int* d_idx = NULL;
	double* d_buf = NULL;   /* rho (N), P/rho^2 (N), 4 per listed particle */
	double* h_out = (double*)malloc((size_t)n_idx * 4u * sizeof(double));
	int rc = 0;
	if (!h_out || hipMalloc((void**)&d_idx, (size_t)n_idx * sizeof(unsigned int)) != hipSuccess ||
	    hipMalloc((void**)&d_buf, (2u * (size_t)N + 4u * (size_t)n_idx) * sizeof(double)) != hipSuccess) {
		sf_error(fn, "allocation failed");
		rc = -1;
		goto done;
	}
	double* rho = d_buf;
	double* cpr = d_buf + N;
	double* out = d_buf + 2u * (size_t)N;
	rc |= sf_hip_check(hipMemcpyAsync(d_idx, idx, (size_t)n_idx * sizeof(unsigned int), hipMemcpyHostToDevice,
	                                  c->stream), fn, "index upload");
	if (!rc)
		rc |= sfd_sph_cubic_f64(sf_blk(c->d_mem, N, SF_E_POS), sf_blk(c->d_mem, N, SF_E_VEL),
		                        sf_blk(c->d_mem, N, SF_E_MAS), sf_blk(c->d_mem, N, SF_E_RAD), c->d_u, N,
		                        p->sph_gamma, d_idx, n_idx, p->acc_mode == SF_ACC_PHYSICAL, rho, cpr, out,
		                        c->stream);
	if (!rc)
		rc |= sf_hip_check(hipMemcpyAsync(h_out, out, (size_t)n_idx * 4u * sizeof(double), hipMemcpyDeviceToHost,
		                                  c->stream), fn, "download");
	if (sf_finish(c, fn)) rc = -1;
	if (!rc) {
		/* rho of the listed particles: one value each, gathered on the host */
		double* h_rho = (double*)malloc((size_t)N * sizeof(double));
		if (!h_rho || sf_hip_check(hipMemcpy(h_rho, rho, (size_t)N * sizeof(double), hipMemcpyDeviceToHost), fn,
		                           "download")) {
			rc = -1;
		} else {
			for (unsigned int k = 0; k < n_idx; ++k) {
				rho_out[k] = h_rho[idx[k]];
				for (int q = 0; q < 3; ++q) acc_out[3u * k + q] = h_out[4u * k + q];
				dudt_out[k] = h_out[4u * k + 3u];
			}
		}
		free(h_rho);
	}
done:
	if (d_idx) {
		hipStreamSynchronize(c->stream);
		hipFree(d_idx);
	}
	if (d_buf) hipFree(d_buf);
	free(h_out);
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
