// adjoint.hip -- vector-Jacobian product of the softened PP force and the
// vector pieces of the adjoint KDK step (no reference counterpart).
//
// Force: a_i = c_i sum_j m_j (x_j - x_i) / s_ij^3, s^2 = |x_i - x_j|^2 + eps^2,
// with c_i = G (physical accelerations) or G m_i (the reference's
// force-like acc).  With d = x_k - x_j and K(d) = I/s^3 - 3 d d^T / s^5,
// the VJP g = (da/dx)^T mu is
//     g_k = G sum_j K(d_kj) (m_k nu_j - m_j nu_k),
// nu = mu (physical) or m mu (reference), since d a_i / d x_i =
// -c_i sum_j m_j K(d_ij) and d a_i / d x_k = c_i m_k K(d_ik).  The pair term
// w = m_k nu_j - m_j nu_k is formed from two rounded products, so the self
// pair (j = k, where K = I/eps^3 is largest) cancels exactly.
//
// Per pair: 26 VALU operations and one v_rsq_f32 (the force takes 13), with
// the same layout as the fast PP kernel: one lane per k, the j stream
// wavefront-uniform (scalar loads), j split into slices over the grid with
// Kahan-folded partial sums per tile, and a fixed-order FP64 combine.  The
// whole tiles of equal-mass particles that sfd_pack_bodies finds form
// w = m (nu_j - nu_k) instead (23 VALU operations; still exactly zero for
// j == k), scaling each tile partial by m.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kTile = 128;
constexpr int kPrefetch = 4;   // j records per scalar-load group
static_assert(kTile % kPrefetch == 0, "tiles hold whole prefetch groups");

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
	const float y = x - c;
	const float t = s + y;
	const float z = t - s;
	c = z - y;
	s = t;
}

// One tile of kTile pairs, j from jt; nbb/nbn hold the tile's first
// kPrefetch j records on entry and the next tile's on exit.  UNIFORM: every
// mass of the tile and of k equals m, and the sums leave it out.
template <bool UNIFORM>
__device__ __forceinline__ void vjp_tile(const float4* __restrict__ body, const float4* __restrict__ nu, unsigned jt,
                                         unsigned j1, float4 (&nbb)[kPrefetch], float4 (&nbn)[kPrefetch],
                                         const float4& bk, const float4& nk, float eps2, float& tx, float& ty,
                                         float& tz) {
	for (unsigned j = jt; j < jt + kTile; j += kPrefetch) {
		float4 cbb[kPrefetch], cbn[kPrefetch];
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) {
			cbb[u] = nbb[u];
			cbn[u] = nbn[u];
		}
		const unsigned jn = (j + kPrefetch < j1) ? j + kPrefetch : j;   // in bounds
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) {
			nbb[u] = body[jn + u];
			nbn[u] = nu[jn + u];
		}
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) {
			const float4 bj = cbb[u];
			const float4 nj = cbn[u];
			const float dx = bk.x - bj.x, dy = bk.y - bj.y, dz = bk.z - bj.z;
			float r2 = __builtin_fmaf(dx, dx, eps2);
			r2 = __builtin_fmaf(dy, dy, r2);
			r2 = __builtin_fmaf(dz, dz, r2);
			const float inv = __builtin_amdgcn_rsqf(r2);
			const float inv2 = inv * inv;
			const float inv3 = inv2 * inv;
			const float i5_3 = inv3 * (3.0f * inv2);
			// exactly zero when j == k: a difference of equal values, or of two
			// equal rounded products
			const float wx = UNIFORM ? nj.x - nk.x : bk.w * nj.x - bj.w * nk.x;
			const float wy = UNIFORM ? nj.y - nk.y : bk.w * nj.y - bj.w * nk.y;
			const float wz = UNIFORM ? nj.z - nk.z : bk.w * nj.z - bj.w * nk.z;
			const float dw = __builtin_fmaf(dz, wz, __builtin_fmaf(dy, wy, dx * wx));
			const float q = i5_3 * dw;
			tx = __builtin_fmaf(-q, dx, __builtin_fmaf(inv3, wx, tx));
			ty = __builtin_fmaf(-q, dy, __builtin_fmaf(inv3, wy, ty));
			tz = __builtin_fmaf(-q, dz, __builtin_fmaf(inv3, wz, tz));
		}
	}
}

// body = (x, y, z, m) and nu = (nu, 0) for n_pad entries (m = nu = 0 past
// N); the word after the n_pad bodies holds n_uni, the leading bodies that
// form whole tiles of one mass, body[0].w (sfd_pack_bodies)
__global__ __launch_bounds__(kBlock) void k_vjp(const float4* __restrict__ body, const float4* __restrict__ nu,
                                                unsigned n_pad, unsigned slices, unsigned slice_len, float eps2,
                                                float4* __restrict__ part) {
	static_assert(kTile == 128u, "n_uni counts whole tiles of 128 bodies");
	const unsigned b = blockIdx.x;
	const unsigned sl = b % slices;
	const unsigned k = (b / slices) * kBlock + threadIdx.x;
	const float4 bk = body[k];
	const float4 nk = nu[k];
	float Sx = 0.0f, Sy = 0.0f, Sz = 0.0f, Cx = 0.0f, Cy = 0.0f, Cz = 0.0f;
	// one group of kPrefetch j records in flight, as in the fast PP kernel
	const unsigned j0 = sl * slice_len;
	const unsigned j1 = min(j0 + slice_len, n_pad);
	const unsigned n_uni = *reinterpret_cast<const unsigned*>(body + n_pad);
	const unsigned ju = min(j1, max(j0, n_uni));
	const float mu = body[0].w;
	float4 nbb[kPrefetch], nbn[kPrefetch];
#pragma unroll
	for (int u = 0; u < kPrefetch; ++u) {
		nbb[u] = body[j0 + u];
		nbn[u] = nu[j0 + u];
	}
	unsigned jt = j0;
	// the equal-mass tiles need m_k = m too: k beyond n_uni (padding, or a
	// partial tile) takes the general sum throughout.  n_uni is a multiple of
	// 128 and a wavefront's 64 k are aligned, so the test is the same on every
	// lane; lane 0's k keeps the branch (and the j stream) scalar.
	if (__builtin_amdgcn_readfirstlane(k) < n_uni) {
		for (; jt < ju; jt += kTile) {
			float tx = 0.0f, ty = 0.0f, tz = 0.0f;
			vjp_tile<true>(body, nu, jt, j1, nbb, nbn, bk, nk, eps2, tx, ty, tz);
			kahan_add(Sx, Cx, mu * tx);
			kahan_add(Sy, Cy, mu * ty);
			kahan_add(Sz, Cz, mu * tz);
		}
	}
	for (; jt < j1; jt += kTile) {
		float tx = 0.0f, ty = 0.0f, tz = 0.0f;
		vjp_tile<false>(body, nu, jt, j1, nbb, nbn, bk, nk, eps2, tx, ty, tz);
		kahan_add(Sx, Cx, tx);
		kahan_add(Sy, Cy, ty);
		kahan_add(Sz, Cz, tz);
	}
	part[(size_t)(2u * sl) * n_pad + k] = make_float4(Sx, Sy, Sz, 0.0f);
	part[(size_t)(2u * sl + 1u) * n_pad + k] = make_float4(Cx, Cy, Cz, 0.0f);
}

// g_k (= or +=) G * sum over slices of (sum - compensation), in FP64
// A block takes kCombI particles; its wavefronts take the slices s = w,
// w + kCombG, ... (w = wavefront) and wavefront 0 adds their sums in order
// (all lanes of one particle per slice would leave most CUs idle at small N).
constexpr unsigned kCombI = 64;
constexpr unsigned kCombG = kBlock / kCombI;
__global__ __launch_bounds__(kBlock) void k_vjp_combine(const float4* __restrict__ part, unsigned n_pad,
                                                        unsigned slices, unsigned N, double G, int accumulate,
                                                        float* __restrict__ g) {
	__shared__ double red[kCombG][3][kCombI];
	const unsigned kl = threadIdx.x % kCombI, w = threadIdx.x / kCombI;
	const unsigned k = blockIdx.x * kCombI + kl;   // < n_pad: the grid covers N in kCombI steps
	double sx = 0.0, sy = 0.0, sz = 0.0;
	for (unsigned s = w; s < slices; s += kCombG) {
		const float4 a = part[(size_t)(2u * s) * n_pad + k];
		const float4 c = part[(size_t)(2u * s + 1u) * n_pad + k];
		sx += (double)a.x - (double)c.x;
		sy += (double)a.y - (double)c.y;
		sz += (double)a.z - (double)c.z;
	}
	red[w][0][kl] = sx;
	red[w][1][kl] = sy;
	red[w][2][kl] = sz;
	__syncthreads();
	if (w != 0 || k >= N) return;
	for (unsigned q = 1; q < kCombG; ++q) {
		sx += red[q][0][kl];
		sy += red[q][1][kl];
		sz += red[q][2][kl];
	}
	float gx = (float)(G * sx), gy = (float)(G * sy), gz = (float)(G * sz);
	if (accumulate) {
		gx += g[3u * k + 0u];
		gy += g[3u * k + 1u];
		gz += g[3u * k + 2u];
	}
	g[3u * k + 0u] = gx;
	g[3u * k + 1u] = gy;
	g[3u * k + 2u] = gz;
}

// nu_i = coef * mu_i (physical) or coef * m_i * mu_i (reference); zero past N
__global__ __launch_bounds__(kBlock) void k_pack_nu(const float* __restrict__ mu, const float* __restrict__ mas,
                                                    unsigned N, unsigned n_pad, float coef, int acc_phys,
                                                    float4* __restrict__ nu) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_pad) return;
	if (i >= N) {
		nu[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		return;
	}
	const float f = acc_phys ? coef : coef * mas[i];
	nu[i] = make_float4(f * mu[3u * i], f * mu[3u * i + 1u], f * mu[3u * i + 2u], 0.0f);
}

// y += a x over n floats
__global__ __launch_bounds__(kBlock) void k_axpy(float* __restrict__ y, const float* __restrict__ x, unsigned n,
                                                 float a) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) y[i] = __builtin_fmaf(a, x[i], y[i]);
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(const float* __restrict__ x, unsigned n,
                                                       double* __restrict__ y) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) y[i] = (double)x[i];
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double* __restrict__ x, unsigned n,
                                                       float* __restrict__ y) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) y[i] = (float)x[i];
}

__global__ __launch_bounds__(kBlock) void k_fill0(float* __restrict__ y, unsigned n) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) y[i] = 0.0f;
}

inline dim3 grid(unsigned n) { return dim3((n + kBlock - 1u) / kBlock); }

}  // namespace

extern "C" {

// g (=, or += when accumulate) G sum_j K(d_kj) (m_k nu_j - m_j nu_k) with
// nu = coef mu (coef m mu): body is the packed (x, m) array of the force
// kernel (n_pad entries), nu and part are scratch (n_pad float4 and
// slices x 2 x n_pad float4).
int sfd_force_vjp(const float* body, const float* mas, const float* mu, unsigned N, unsigned n_pad, float eps2,
                  double G, int acc_phys, float coef, float* nu, float* part, unsigned slices, int accumulate,
                  float* g, hipStream_t s) {
	if (N == 0) return 0;
	if (n_pad % kBlock != 0 || slices == 0) return -1;
	float4* nu4 = reinterpret_cast<float4*>(nu);
	hipLaunchKernelGGL(k_pack_nu, grid(n_pad), dim3(kBlock), 0, s, mu, mas, N, n_pad, coef, acc_phys, nu4);
	// whole tiles per slice, every slice non-empty (at most the requested count)
	const unsigned tiles = n_pad / kTile;
	if (slices > tiles) slices = tiles;
	const unsigned slice_len = (tiles + slices - 1u) / slices * kTile;
	slices = (n_pad + slice_len - 1u) / slice_len;
	const unsigned grid_v = (n_pad / kBlock) * slices;
	const int tok = sf_prof_begin("force_vjp", s);
	hipLaunchKernelGGL(k_vjp, dim3(grid_v), dim3(kBlock), 0, s, reinterpret_cast<const float4*>(body), nu4, n_pad,
	                   slices, slice_len, eps2, reinterpret_cast<float4*>(part));
	sf_prof_end(tok, s);
	hipLaunchKernelGGL(k_vjp_combine, dim3((N + kCombI - 1) / kCombI), dim3(kBlock), 0, s,
	                   reinterpret_cast<const float4*>(part), n_pad, slices, N, G, accumulate, g);
	return sf_hip_check(hipGetLastError(), "sfd_force_vjp", "launch");
}

int sfd_axpy(float* y, const float* x, unsigned n, float a, hipStream_t s) {
	if (n == 0) return 0;
	hipLaunchKernelGGL(k_axpy, grid(n), dim3(kBlock), 0, s, y, x, n, a);
	return sf_hip_check(hipGetLastError(), "sfd_axpy", "launch");
}

int sfd_f32_to_f64(const float* x, unsigned n, double* y, hipStream_t s) {
	if (n == 0) return 0;
	hipLaunchKernelGGL(k_f32_to_f64, grid(n), dim3(kBlock), 0, s, x, n, y);
	return sf_hip_check(hipGetLastError(), "sfd_f32_to_f64", "launch");
}

int sfd_f64_to_f32(const double* x, unsigned n, float* y, hipStream_t s) {
	if (n == 0) return 0;
	hipLaunchKernelGGL(k_f64_to_f32, grid(n), dim3(kBlock), 0, s, x, n, y);
	return sf_hip_check(hipGetLastError(), "sfd_f64_to_f32", "launch");
}

int sfd_fill0(float* y, unsigned n, hipStream_t s) {
	if (n == 0) return 0;
	hipLaunchKernelGGL(k_fill0, grid(n), dim3(kBlock), 0, s, y, n);
	return sf_hip_check(hipGetLastError(), "sfd_fill0", "launch");
}

}  // extern "C"
