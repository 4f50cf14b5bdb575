/*
 * gravity.c -- solve_gravity_part_part with the reference's signature
 * (src/gravity.h:22, src/gravity.c:666-819), the FP64 direct-sum precision
 * reference, and the device-array force entry used by the adjoint.
 */
#include <stdio.h>
#include <stdlib.h>

#include "sf_host.h"

/* j slices of the fast kernel: enough blocks to give every CU several waves */
static unsigned int pp_slices(unsigned int n_pad) {
	/* one j range per wavefront of 64 i; about 128k wavefronts (measured best
	   at N = 262144: 32 ranges) while each range keeps at least 256 j */
	const unsigned int iblocks = n_pad / 256u;
	unsigned int s = iblocks ? (32768u + iblocks - 1u) / iblocks : 1u;
	const unsigned int smax = n_pad / 256u;
	if (s > smax) s = smax;
	if (s > 64u) s = 64u;
	if (s < 1u) s = 1u;
	return s;
}

/* reference: (real)(EPSILON*EPSILON) is squared in double, then rounded */
static float eps2_f32(double eps) { return (float)(eps * eps); }

static int pp_on(const sf_params_t* p, unsigned int N, unsigned int n_pad, const float* pos, const float* mas,
                 float* body, float** part, unsigned int* part_slices, sf_ctx* c, float* acc, float* pot,
                 hipStream_t s) {
	const int phys = p->acc_mode == SF_ACC_PHYSICAL;
	if (p->precision == SF_PRECISION_EXACT) {
		return sfd_pp_exact(pos, mas, N, eps2_f32(p->epsilon), (float)(-p->G), phys, p->kahan_solver,
		                    p->kahan_energy, acc, pot, s);
	}
	const unsigned int slices = pp_slices(n_pad);
	if (c) {
		if (sf_ctx_reserve_part(c, slices)) return -1;
		*part = c->d_part;
	} else if (*part_slices < slices) {
		return -1;
	}
	if (sfd_pack_bodies(pos, mas, N, n_pad, body, s)) return -1;
	return sfd_pp_fast(body, N, n_pad, eps2_f32(p->epsilon), p->G, phys, *part, slices, acc, pot, s);
}

unsigned int sf_pp_slice_count(unsigned int n_pad) { return pp_slices(n_pad); }

size_t sf_pp_part_bytes(unsigned int n_pad) { return sfd_pp_part_bytes(n_pad, pp_slices(n_pad)); }

int sf_pp_force_arrays(const sf_params_t* p, unsigned int N, unsigned int n_pad, const float* pos, const float* mas,
                       float* body, float* part, size_t part_bytes, float* acc, float* pot, hipStream_t s) {
	if (part_bytes < sf_pp_part_bytes(n_pad)) {
		sf_error("sf_pp_force_arrays", "%zu bytes of slice partials, %zu needed", part_bytes, sf_pp_part_bytes(n_pad));
		return -1;
	}
	if (sfd_pack_bodies(pos, mas, N, n_pad, body, s)) return -1;
	return sfd_pp_fast(body, N, n_pad, eps2_f32(p->epsilon), p->G, p->acc_mode == SF_ACC_PHYSICAL, part,
	                   pp_slices(n_pad), acc, pot, s);
}

int sf_solve_gravity_pp(sf_ctx* c) {
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	float* part = NULL;
	unsigned int ps = 0;
	return pp_on(p, N, c->n_pad, sf_blk(c->d_mem, N, SF_E_POS), sf_blk(c->d_mem, N, SF_E_MAS), c->d_body, &part, &ps,
	             c, sf_blk(c->d_mem, N, SF_E_ACC), sf_blk(c->d_mem, N, SF_E_POT), c->stream);
}

/* reference src/gravity.c:667-819: reads pos, mas; writes acc (assigned, not
   accumulated) and pot. */
int solve_gravity_part_part(const sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "solve_gravity_part_part");
	if (!c) return STARFLOOD_FAILURE;
	if (sf_upload_blocks(c, SF_E_POS, 3u) || sf_upload_blocks(c, SF_E_MAS, 1u)) return STARFLOOD_FAILURE;
	int rc = sf_solve_gravity_pp(c);
	if (sf_download_blocks(c, SF_E_ACC, 3u) || sf_download_blocks(c, SF_E_POT, 1u)) rc = -1;
	if (sf_finish(c, "solve_gravity_part_part")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

int sf_pp_direct_f64(const sim_t* sim_ptr, const unsigned int* idx, unsigned int n_idx, double* acc_out,
                     double* pot_out) {
	if (!sim_ptr || (!idx && n_idx) || (n_idx && (!acc_out || !pot_out))) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_pp_direct_f64");
	if (!c) return STARFLOOD_FAILURE;
	if (n_idx == 0) return STARFLOOD_SUCCESS;
	for (unsigned int k = 0; k < n_idx; ++k) {
		if (idx[k] >= c->N) {
			sf_error("sf_pp_direct_f64", "index %u out of range (N = %u)", idx[k], c->N);
			return STARFLOOD_FAILURE;
		}
	}
	const sf_params_t* p = sf_params();
	if (sf_upload_blocks(c, SF_E_POS, 3u) || sf_upload_blocks(c, SF_E_MAS, 1u)) return STARFLOOD_FAILURE;
	unsigned int* d_idx = NULL;
	double* d_out = NULL;
	int rc = 0;
	if (hipMalloc((void**)&d_idx, (size_t)n_idx * sizeof(unsigned int)) != hipSuccess ||
	    hipMalloc((void**)&d_out, (size_t)n_idx * 4u * sizeof(double)) != hipSuccess) {
		sf_error("sf_pp_direct_f64", "device allocation failed");
		rc = -1;
		goto done;
	}
	if (sf_hip_check(hipMemcpyAsync(d_idx, idx, (size_t)n_idx * sizeof(unsigned int), hipMemcpyHostToDevice,
	                                c->stream),
	                 "sf_pp_direct_f64", "index upload")) {
		rc = -1;
		goto done;
	}
	if (sfd_pp_f64(sf_blk(c->d_mem, c->N, SF_E_POS), sf_blk(c->d_mem, c->N, SF_E_MAS), c->N, p->epsilon * p->epsilon,
	               p->G, p->acc_mode == SF_ACC_PHYSICAL, d_idx, n_idx, d_out, d_out + 3u * (size_t)n_idx,
	               c->stream)) {
		rc = -1;
		goto done;
	}
	if (sf_hip_check(hipMemcpyAsync(acc_out, d_out, (size_t)n_idx * 3u * sizeof(double), hipMemcpyDeviceToHost,
	                                c->stream),
	                 "sf_pp_direct_f64", "download") ||
	    sf_hip_check(hipMemcpyAsync(pot_out, d_out + 3u * (size_t)n_idx, (size_t)n_idx * sizeof(double),
	                                hipMemcpyDeviceToHost, c->stream),
	                 "sf_pp_direct_f64", "download"))
		rc = -1;
	if (sf_finish(c, "sf_pp_direct_f64")) rc = -1;
done:
	if (d_idx) {
		hipStreamSynchronize(c->stream);
		hipFree(d_idx);
	}
	if (d_out) hipFree(d_out);
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
