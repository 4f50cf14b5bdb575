// sph_f64.hip -- the cubic-spline ideal-gas SPH model in FP64, every pair,
// for a list of particles: the precision reference of the FP32 SPH passes
// (as k_pp_f64 is of the PP kernel), not a solver.  No counterpart in the
// reference.
//
//   rho_i = sum_j m_j 0.5 (W(r, h_i) + W(r, h_j)),  P_i = (gamma - 1) rho_i u_i
//   a_i   = -s_i sum_{j != i} m_j (P_i/rho_i^2 + P_j/rho_j^2) dWbar/dr r_ij / r
//   du_i  = 0.5 sum_{j != i} m_j (P_i/rho_i^2 + P_j/rho_j^2) dWbar/dr (v_ij . r_ij) / r
//
// with W the M4 spline of support radius h and s_i = 1 (physical) or m_i
// (the reference's force-like acc).  One block per particle, lanes stride
// over j, and a fixed-shape tree reduction in LDS, so the result does not
// depend on scheduling.  The force rows need every neighbour's P/rho^2, so
// the density pass runs over all N.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double w_cubic(double r, double h) {
	const double q = r / h, k = 8.0 / (kPi * h * h * h);
	if (q <= 0.5) return k * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
	if (q <= 1.0) return k * 2.0 * (1.0 - q) * (1.0 - q) * (1.0 - q);
	return 0.0;
}

__device__ __forceinline__ double dw_cubic(double r, double h) {
	const double q = r / h, k = 8.0 / (kPi * h * h * h * h);
	if (q <= 0.5) return k * (-12.0 * q + 18.0 * q * q);
	if (q <= 1.0) return k * -6.0 * (1.0 - q) * (1.0 - q);
	return 0.0;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*red)[kBlock]) {
#pragma unroll
	for (int q = 0; q < NV; ++q) red[q][threadIdx.x] = v[q];
	__syncthreads();
	for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
#pragma unroll
			for (int q = 0; q < NV; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
		}
		__syncthreads();
	}
#pragma unroll
	for (int q = 0; q < NV; ++q) v[q] = red[q][0];
}

// rho_i and c_i = P_i / rho_i^2 of every particle
__global__ __launch_bounds__(kBlock) void k_sph_rho_f64(const float* __restrict__ pos, const float* __restrict__ mas,
                                                        const float* __restrict__ rad, const float* __restrict__ u,
                                                        unsigned N, double gamma, double* __restrict__ rho,
                                                        double* __restrict__ cpr) {
	__shared__ double red[1][kBlock];
	const unsigned i = blockIdx.x;
	const double xi = pos[3u * i], yi = pos[3u * i + 1u], zi = pos[3u * i + 2u], hi = rad[i];
	double s[1] = {0.0};
	for (unsigned j = threadIdx.x; j < N; j += kBlock) {
		const double dx = xi - (double)pos[3u * j], dy = yi - (double)pos[3u * j + 1u],
		             dz = zi - (double)pos[3u * j + 2u];
		const double r2 = dx * dx + dy * dy + dz * dz, hj = rad[j];
		if (r2 > hi * hi && r2 > hj * hj) continue;   // outside both supports: W = 0
		const double r = sqrt(r2);
		s[0] += (double)mas[j] * 0.5 * (w_cubic(r, hi) + w_cubic(r, hj));
	}
	block_sum<1>(s, red);
	if (threadIdx.x == 0) {
		rho[i] = s[0];
		cpr[i] = s[0] > 0.0 ? (gamma - 1.0) * (double)u[i] / s[0] : 0.0;   // (gamma-1) rho u / rho^2
	}
}

__global__ __launch_bounds__(kBlock) void k_sph_force_f64(const float* __restrict__ pos, const float* __restrict__ vel,
                                                          const float* __restrict__ mas, const float* __restrict__ rad,
                                                          const double* __restrict__ cpr, unsigned N,
                                                          const unsigned* __restrict__ idx, int acc_phys,
                                                          double* __restrict__ out) {
	__shared__ double red[4][kBlock];
	const unsigned i = idx[blockIdx.x];
	const double xi = pos[3u * i], yi = pos[3u * i + 1u], zi = pos[3u * i + 2u], hi = rad[i];
	const double vxi = vel[3u * i], vyi = vel[3u * i + 1u], vzi = vel[3u * i + 2u], ci = cpr[i];
	double s[4] = {0.0, 0.0, 0.0, 0.0};
	for (unsigned j = threadIdx.x; j < N; j += kBlock) {
		if (j == i) continue;
		const double dx = xi - (double)pos[3u * j], dy = yi - (double)pos[3u * j + 1u],
		             dz = zi - (double)pos[3u * j + 2u];
		const double r2 = dx * dx + dy * dy + dz * dz, hj = rad[j];
		if (r2 <= 0.0 || (r2 > hi * hi && r2 > hj * hj)) continue;
		const double r = sqrt(r2);
		const double f = (double)mas[j] * (ci + cpr[j]) * 0.5 * (dw_cubic(r, hi) + dw_cubic(r, hj)) / r;
		const double vd = (vxi - (double)vel[3u * j]) * dx + (vyi - (double)vel[3u * j + 1u]) * dy +
		                  (vzi - (double)vel[3u * j + 2u]) * dz;
		s[0] -= f * dx;
		s[1] -= f * dy;
		s[2] -= f * dz;
		s[3] += 0.5 * f * vd;
	}
	block_sum<4>(s, red);
	if (threadIdx.x == 0) {
		const double sc = acc_phys ? 1.0 : (double)mas[i];
		out[4u * blockIdx.x + 0u] = sc * s[0];
		out[4u * blockIdx.x + 1u] = sc * s[1];
		out[4u * blockIdx.x + 2u] = sc * s[2];
		out[4u * blockIdx.x + 3u] = s[3];
	}
}

}  // namespace

// rho (N), c (N) scratch; out: 4 doubles per listed particle (acc xyz, du/dt)
extern "C" int sfd_sph_cubic_f64(const float* pos, const float* vel, const float* mas, const float* rad,
                                 const float* u, unsigned N, double gamma, const unsigned* idx, unsigned n_idx,
                                 int acc_phys, double* rho, double* cpr, double* out, hipStream_t s) {
	if (N == 0) return 0;
	const int tok = sf_prof_begin("sph_f64", s);
	hipLaunchKernelGGL(k_sph_rho_f64, dim3(N), dim3(kBlock), 0, s, pos, mas, rad, u, N, gamma, rho, cpr);
	if (n_idx)
		hipLaunchKernelGGL(k_sph_force_f64, dim3(n_idx), dim3(kBlock), 0, s, pos, vel, mas, rad, cpr, N, idx,
		                   acc_phys, out);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_sph_cubic_f64", "launch");
}
