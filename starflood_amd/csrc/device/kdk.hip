// kdk.hip -- kick-drift-kick pieces, per-particle energies and FP64
// reductions of the state (reference src/simulation.c:381-504).
//
// Exact policy: FP32 elementwise updates in the reference's operation order
// (file compiled with -ffp-contract=off) and a sequential FP64 Kahan sum in
// index order, so T and V are bit-identical to the reference's.
// Fast policy: kick and FP64 drift fused in one pass over the state, and
// tree-shaped FP64 block reductions with a fixed grid, which are
// deterministic and more accurate than a sequential sum.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kRedBlocks = 512;   // fixed grid of the state reduction

inline unsigned grid_for(unsigned n) { return (n + kBlock - 1) / kBlock; }

// vel[k] += (0.5 dt) acc[k]  (src/simulation.c:382, :424)
__global__ __launch_bounds__(kBlock) void k_kick_exact(float* __restrict__ vel, const float* __restrict__ acc,
                                                       unsigned n3, float half_dt) {
	const unsigned k = blockIdx.x * kBlock + threadIdx.x;
	if (k < n3) vel[k] = vel[k] + half_dt * acc[k];
}

// pos[k] += dt vel[k]  (src/simulation.c:394)
__global__ __launch_bounds__(kBlock) void k_drift_exact(float* __restrict__ pos, const float* __restrict__ vel,
                                                        unsigned n3, float dt) {
	const unsigned k = blockIdx.x * kBlock + threadIdx.x;
	if (k < n3) pos[k] = pos[k] + dt * vel[k];
}

// ken_i = (0.5 m_i) * ((vx^2 + vy^2) + vz^2); pen_i = m_i pot_i  (src/simulation.c:434-451)
__global__ __launch_bounds__(kBlock) void k_energy_terms(const float* __restrict__ vel, const float* __restrict__ mas,
                                                         const float* __restrict__ pot, unsigned N,
                                                         float* __restrict__ ken, float* __restrict__ pen) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	const float m = mas[i];
	const float vx = vel[3u * i], vy = vel[3u * i + 1u], vz = vel[3u * i + 2u];
	const float v2 = (vx * vx + vy * vy) + vz * vz;
	ken[i] = 0.5f * m * v2;
	pen[i] = m * pot[i];
}

// Sequential FP64 Kahan in index order; a single lane, for bit parity with
// src/simulation.c:455-497.
__global__ void k_energy_sum_sequential(const float* __restrict__ x, unsigned N, double* __restrict__ out) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	double S = 0.0, C = 0.0;
	for (unsigned i = 0; i < N; ++i) {
		const double y = (double)x[i] - C;
		const double t = S + y;
		const double z = t - S;
		C = z - y;
		S = t;
	}
	*out = S;
}

// x64 follows pos unless it rounds to it: a host-side change of pos (or a
// first use) restarts the FP64 track from the FP32 value.
__global__ __launch_bounds__(kBlock) void k_x64_refresh(const float* __restrict__ pos, double* __restrict__ x64,
                                                        unsigned n3) {
	const unsigned k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= n3) return;
	const float p = pos[k];
	if ((float)x64[k] != p) x64[k] = (double)p;
}

// refresh: first apply k_x64_refresh's rule to x64 (a caller that always
// writes pos from x64 passes 0)
__global__ __launch_bounds__(kBlock) void k_kick_drift_fast(float* __restrict__ vel, const float* __restrict__ acc,
                                                            double* __restrict__ x64, float* __restrict__ pos,
                                                            unsigned n3, float half_dt, double dt, int refresh) {
	const unsigned k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= n3) return;
	const float v = __builtin_fmaf(half_dt, acc[k], vel[k]);
	vel[k] = v;
	double x0 = x64[k];
	if (refresh) {
		const float p = pos[k];
		if ((float)x0 != p) x0 = (double)p;
	}
	const double x = __builtin_fma(dt, (double)v, x0);
	x64[k] = x;
	pos[k] = (float)x;
}

__global__ __launch_bounds__(kBlock) void k_kick_fast(float* __restrict__ vel, const float* __restrict__ acc,
                                                      unsigned n3, float half_dt) {
	const unsigned k = blockIdx.x * kBlock + threadIdx.x;
	if (k < n3) vel[k] = __builtin_fmaf(half_dt, acc[k], vel[k]);
}

// k_kick_fast followed by k_energy_terms, per particle, with the same
// operations: one pass over vel instead of two launches.
__global__ __launch_bounds__(kBlock) void k_kick_energy_fast(float* __restrict__ vel, const float* __restrict__ acc,
                                                             const float* __restrict__ mas,
                                                             const float* __restrict__ pot, unsigned N, float half_dt,
                                                             float* __restrict__ ken, float* __restrict__ pen) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	const float vx = __builtin_fmaf(half_dt, acc[3u * i], vel[3u * i]);
	const float vy = __builtin_fmaf(half_dt, acc[3u * i + 1u], vel[3u * i + 1u]);
	const float vz = __builtin_fmaf(half_dt, acc[3u * i + 2u], vel[3u * i + 2u]);
	vel[3u * i] = vx;
	vel[3u * i + 1u] = vy;
	vel[3u * i + 2u] = vz;
	const float m = mas[i];
	const float v2 = (vx * vx + vy * vy) + vz * vz;
	ken[i] = 0.5f * m * v2;
	pen[i] = m * pot[i];
}

constexpr int kRedQ = 11;   // T, V, W, U, Px, Py, Pz, Lx, Ly, Lz, K

// Per-block partial sums with a fixed grid-stride assignment and a fixed
// LDS tree, so the result does not depend on scheduling.
__global__ __launch_bounds__(kBlock) void k_reduce_state(const float* __restrict__ pos, const double* __restrict__ x64,
                                                         const float* __restrict__ vel, const float* __restrict__ mas,
                                                         const float* __restrict__ pot, const float* __restrict__ ken,
                                                         const float* __restrict__ pen, const float* __restrict__ u,
                                                         unsigned N, double G, double eps, double* __restrict__ red) {
	__shared__ double sh[kRedQ][kBlock];
	double q[kRedQ];
#pragma unroll
	for (int t = 0; t < kRedQ; ++t) q[t] = 0.0;
	const double self = eps > 0.0 ? G / eps : 0.0;
	for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
		const double m = mas[i];
		double x, y, z;
		if (x64) {
			x = x64[3u * i]; y = x64[3u * i + 1u]; z = x64[3u * i + 2u];
		} else {
			x = pos[3u * i]; y = pos[3u * i + 1u]; z = pos[3u * i + 2u];
		}
		const double vx = vel[3u * i], vy = vel[3u * i + 1u], vz = vel[3u * i + 2u];
		q[0] += (double)ken[i];
		q[1] += (double)pen[i];
		q[2] += 0.5 * m * ((double)pot[i] + self * m);
		if (u) q[3] += m * (double)u[i];
		q[4] += m * vx;
		q[5] += m * vy;
		q[6] += m * vz;
		q[7] += m * (y * vz - z * vy);
		q[8] += m * (z * vx - x * vz);
		q[9] += m * (x * vy - y * vx);
		q[10] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
	}
#pragma unroll
	for (int t = 0; t < kRedQ; ++t) sh[t][threadIdx.x] = q[t];
	__syncthreads();
	for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
#pragma unroll
			for (int t = 0; t < kRedQ; ++t) sh[t][threadIdx.x] += sh[t][threadIdx.x + h];
		}
		__syncthreads();
	}
	if (threadIdx.x < kRedQ) red[(size_t)blockIdx.x * SF_RED_VALUES + threadIdx.x] = sh[threadIdx.x][0];
}

// Second stage: one block folds the block partials in a fixed tree and
// writes the totals to red[0 .. SF_RED_VALUES) (after the partials region).
__global__ __launch_bounds__(kBlock) void k_reduce_final(double* __restrict__ red, unsigned nblocks,
                                                         double* __restrict__ out) {
	__shared__ double sh[kRedQ][kBlock];
	double q[kRedQ];
#pragma unroll
	for (int t = 0; t < kRedQ; ++t) q[t] = 0.0;
	for (unsigned b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
		for (int t = 0; t < kRedQ; ++t) q[t] += red[(size_t)b * SF_RED_VALUES + t];
	}
#pragma unroll
	for (int t = 0; t < kRedQ; ++t) sh[t][threadIdx.x] = q[t];
	__syncthreads();
	for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
#pragma unroll
			for (int t = 0; t < kRedQ; ++t) sh[t][threadIdx.x] += sh[t][threadIdx.x + h];
		}
		__syncthreads();
	}
	if (threadIdx.x < (unsigned)kRedQ) out[threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace

extern "C" int sfd_kick_exact(float* vel, const float* acc, unsigned n3, float half_dt, hipStream_t s) {
	if (!n3) return 0;
	const int tok = sf_prof_begin("kick", s);
	hipLaunchKernelGGL(k_kick_exact, dim3(grid_for(n3)), dim3(kBlock), 0, s, vel, acc, n3, half_dt);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_kick_exact", "launch");
}

extern "C" int sfd_drift_exact(float* pos, const float* vel, unsigned n3, float dt, hipStream_t s) {
	if (!n3) return 0;
	const int tok = sf_prof_begin("drift", s);
	hipLaunchKernelGGL(k_drift_exact, dim3(grid_for(n3)), dim3(kBlock), 0, s, pos, vel, n3, dt);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_drift_exact", "launch");
}

extern "C" int sfd_energy_terms(const float* vel, const float* mas, const float* pot, unsigned N, float* ken,
                                float* pen, hipStream_t s) {
	if (!N) return 0;
	const int tok = sf_prof_begin("energy_terms", s);
	hipLaunchKernelGGL(k_energy_terms, dim3(grid_for(N)), dim3(kBlock), 0, s, vel, mas, pot, N, ken, pen);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_energy_terms", "launch");
}

extern "C" int sfd_energy_sum_sequential(const float* x, unsigned N, double* out, hipStream_t s) {
	const int tok = sf_prof_begin("energy_sum_seq", s);
	hipLaunchKernelGGL(k_energy_sum_sequential, dim3(1), dim3(64), 0, s, x, N, out);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_energy_sum_sequential", "launch");
}

extern "C" int sfd_x64_refresh(const float* pos, double* x64, unsigned n3, hipStream_t s) {
	if (!n3) return 0;
	hipLaunchKernelGGL(k_x64_refresh, dim3(grid_for(n3)), dim3(kBlock), 0, s, pos, x64, n3);
	return sf_hip_check(hipGetLastError(), "sfd_x64_refresh", "launch");
}

extern "C" int sfd_kick_drift_fast(float* vel, const float* acc, double* x64, float* pos, unsigned n3,
                                   float half_dt, double dt, int refresh, hipStream_t s) {
	if (!n3) return 0;
	const int tok = sf_prof_begin("kick_drift", s);
	hipLaunchKernelGGL(k_kick_drift_fast, dim3(grid_for(n3)), dim3(kBlock), 0, s, vel, acc, x64, pos, n3, half_dt,
	                   dt, refresh);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_kick_drift_fast", "launch");
}

extern "C" int sfd_kick_energy_fast(float* vel, const float* acc, const float* mas, const float* pot, unsigned N,
                                    float half_dt, float* ken, float* pen, hipStream_t s) {
	if (!N) return 0;
	const int tok = sf_prof_begin("kick", s);
	hipLaunchKernelGGL(k_kick_energy_fast, dim3(grid_for(N)), dim3(kBlock), 0, s, vel, acc, mas, pot, N, half_dt, ken,
	                   pen);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_kick_energy_fast", "launch");
}

extern "C" int sfd_kick_fast(float* vel, const float* acc, unsigned n3, float half_dt, hipStream_t s) {
	if (!n3) return 0;
	const int tok = sf_prof_begin("kick", s);
	hipLaunchKernelGGL(k_kick_fast, dim3(grid_for(n3)), dim3(kBlock), 0, s, vel, acc, n3, half_dt);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_kick_fast", "launch");
}

extern "C" unsigned sfd_reduce_scratch_doubles(unsigned N) {
	(void)N;
	return (kRedBlocks + 1u) * SF_RED_VALUES;
}

extern "C" int sfd_reduce_state(const float* pos, const double* x64, const float* vel, const float* mas,
                                const float* pot, const float* ken, const float* pen, const float* u, unsigned N,
                                double G, double eps, double* red, hipStream_t s) {
	unsigned nb = grid_for(N);
	if (nb > kRedBlocks) nb = kRedBlocks;
	if (nb == 0) nb = 1;
	const int tok = sf_prof_begin("reduce_state", s);
	hipLaunchKernelGGL(k_reduce_state, dim3(nb), dim3(kBlock), 0, s, pos, x64, vel, mas, pot, ken, pen, u, N, G, eps,
	                   red);
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, s, red, nb, red + (size_t)kRedBlocks * SF_RED_VALUES);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_reduce_state", "launch");
}
