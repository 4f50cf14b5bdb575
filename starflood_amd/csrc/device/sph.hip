// sph.hip -- smoothed-particle hydrodynamics sums, all pairs (PP SPH).
//
// SF_SPH_REFERENCE: the reference's model (src/sph.c:14-283) in its
// operation order -- Gaussian kernel W = (1/(sqrt(pi) h))^3 exp(-r^2/h^2),
// rho_i = sum_j 0.5 (W_i + W_j) m_j over every j, polytrope P = (0.1 rho)^2,
// and acc_i += -m_i sum_{r^2 <= 1} (P_i/rho_i^2 + P_j/rho_j^2) 0.5 (n_i + n_j) r_ij
// with n = -2 exp(-r^2/h^2) / (pi^{3/2} h^5).  The file is compiled with
// -ffp-contract=off and IEEE division and square root, so every product,
// quotient and Kahan step rounds as the reference's does; the per-particle
// factors that do not depend on the pair ((a^2) a, h h, pi^{3/2} h^5 and
// P/rho^2) are computed once per particle with the same operations, which
// gives the same bits.  exp() is the device library's (faithfully rounded,
// like glibc's expf, but not always the same float) and (0.1 rho)^2 is the
// correctly rounded product where glibc's powf is off by one ulp in about
// 0.07% of cases: results agree with the reference to a few ulp, not bit
// for bit.
//
// SF_SPH_CUBIC_IDEAL: cubic spline W(r, h) = 8/(pi h^3) {1 - 6q^2 + 6q^3,
// q <= 1/2; 2 (1-q)^3, q <= 1} (support radius h, GADGET-2 convention), the
// symmetrised kernel 0.5 (W(r, h_i) + W(r, h_j)), ideal gas P = (gamma-1) rho u,
// acc_i += -s_i sum_j m_j (P_i/rho_i^2 + P_j/rho_j^2) grad_i W and
// du_i/dt = 0.5 sum_j m_j (P_i/rho_i^2 + P_j/rho_j^2) v_ij . grad_i W, where
// s_i = 1 for physical accelerations and m_i for the reference's force-like
// acc.  The j stream of packed records is wavefront-uniform, so the compiler
// reads it through scalar loads; lanes with no neighbour skip the kernel
// arithmetic.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kTj = 128;   // j per unrolled tile of the cubic-model passes

// the reference's constants (src/common.h:18, :29), rounded to float as
// (real)INV_SQRT_PI and (real)INV_POW_PI_3_2 are; the second one's value
// is pi^{3/2}
constexpr float kInvSqrtPi = (float)0.56418958354775628694807945156077258584405062932900;
constexpr float kPiPow32 = (float)5.5683279968317078452848179821188357020136243902832;
constexpr float kPi = 3.14159265358979323846f;

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
	const float y = x - c;
	const float t = s + y;
	const float z = t - s;
	c = z - y;
	s = t;
}

// ------------------------------------------------------------ reference model
// a3 = (a a) a with a = INV_SQRT_PI / h (src/sph.c:116-119); hh = h h;
// den = INV_POW_PI_3_2 h h h h h evaluated left to right (src/sph.c:239-240)
__global__ __launch_bounds__(kBlock) void k_tables_ref(const float* __restrict__ rad, unsigned N,
                                                       float* __restrict__ a3, float* __restrict__ hh,
                                                       float* __restrict__ den) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	const float h = rad[i];
	const float a = kInvSqrtPi / h;
	a3[i] = a * a * a;
	hh[i] = h * h;
	den[i] = kPiPow32 * h * h * h * h * h;
}

// src/sph.c:35-145
template <bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_rho_ref(const float* __restrict__ pos, const float* __restrict__ mas,
                                                    const float* __restrict__ a3, const float* __restrict__ hh,
                                                    unsigned N, float* __restrict__ rho) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	const unsigned ii = i < N ? i : N - 1u;
	const float xi = pos[3u * ii], yi = pos[3u * ii + 1u], zi = pos[3u * ii + 2u];
	const float a3i = a3[ii], hhi = hh[ii];
	float s = 0.0f, c = 0.0f;
	for (unsigned j = 0u; j < N; ++j) {
		const float mj = mas[j];
		const float rx = xi - pos[3u * j], ry = yi - pos[3u * j + 1u], rz = zi - pos[3u * j + 2u];
		const float r2 = (rx * rx) + (ry * ry) + (rz * rz);
		const float wi = a3i * expf(-r2 / hhi);
		const float wj = a3[j] * expf(-r2 / hh[j]);
		const float t = 0.5f * (wi + wj) * mj;
		if (KAHAN) kahan_add(s, c, t);
		else s += t;
	}
	if (i < N) rho[i] = s;
}

// P = (K rho)^(1/n + 1) with K = 0.1, n = 1 (src/sph.c:147-157), and the
// momentum-equation factor P / rho^2 (src/sph.c:189-192, :225-226)
__global__ __launch_bounds__(kBlock) void k_eos_ref(const float* __restrict__ rho, unsigned N,
                                                    float* __restrict__ prs, float* __restrict__ con) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	const float x = 0.100f * rho[i];
	const float p = x * x;
	prs[i] = p;
	const float r2 = rho[i] * rho[i];
	con[i] = 0.0f < r2 ? p / r2 : 0.0f;
}

// src/sph.c:163-276
template <bool KAHAN>
__global__ __launch_bounds__(kBlock) void k_force_ref(const float* __restrict__ pos, const float* __restrict__ mas,
                                                      const float* __restrict__ hh, const float* __restrict__ den,
                                                      const float* __restrict__ con, unsigned N,
                                                      float* __restrict__ acc) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	const unsigned ii = i < N ? i : N - 1u;
	const float xi = pos[3u * ii], yi = pos[3u * ii + 1u], zi = pos[3u * ii + 2u];
	const float hhi = hh[ii], deni = den[ii], coni = con[ii];
	float sx = 0.0f, sy = 0.0f, sz = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
	for (unsigned j = 0u; j < N; ++j) {
		const float rx = xi - pos[3u * j], ry = yi - pos[3u * j + 1u], rz = zi - pos[3u * j + 2u];
		const float r2 = (rx * rx) + (ry * ry) + (rz * rz);
		if (r2 > 1.0f) continue;
		const float ni = -2.0f * expf(-r2 / hhi) / deni;
		const float nj = -2.0f * expf(-r2 / hh[j]) / den[j];
		const float g = 0.5f * (ni + nj);
		const float cc = coni + con[j];
		const float X = cc * (g * rx), Y = cc * (g * ry), Z = cc * (g * rz);
		if (KAHAN) {
			kahan_add(sx, cx, X);
			kahan_add(sy, cy, Y);
			kahan_add(sz, cz, Z);
		} else {
			sx += X;
			sy += Y;
			sz += Z;
		}
	}
	if (i < N) {
		const float nm = -mas[i];
		acc[3u * i + 0u] += nm * sx;
		acc[3u * i + 1u] += nm * sy;
		acc[3u * i + 2u] += nm * sz;
	}
}

// ---------------------------------------------------------------- cubic model
// Packed per-particle records (n_pad entries, zero past N):
//   q0 = (x, y, z, h^2), q1 = (m, 1/h, u_pred, P/rho^2), q2 = (v_pred, 0).
// v_pred = v + pred_dt acc and u_pred = u + pred_dt du/dt predict the
// half-kicked state to the end of the step (pred_dt = 0 takes it as it is).
__global__ __launch_bounds__(kBlock) void k_pack_cubic(const float* __restrict__ pos, const float* __restrict__ vel,
                                                       const float* __restrict__ acc, const float* __restrict__ mas,
                                                       const float* __restrict__ rad, const float* __restrict__ u,
                                                       const float* __restrict__ dudt, unsigned N, unsigned n_pad,
                                                       float pred_dt, float4* __restrict__ q0,
                                                       float4* __restrict__ q1, float4* __restrict__ q2) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_pad) return;
	if (i >= N) {
		const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		q0[i] = z;
		q1[i] = z;
		q2[i] = z;
		return;
	}
	const float h = rad[i];
	q0[i] = make_float4(pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u], h * h);
	q1[i] = make_float4(mas[i], 1.0f / h, __builtin_fmaf(pred_dt, dudt[i], u[i]), 0.0f);
	q2[i] = make_float4(__builtin_fmaf(pred_dt, acc[3u * i], vel[3u * i]),
	                    __builtin_fmaf(pred_dt, acc[3u * i + 1u], vel[3u * i + 1u]),
	                    __builtin_fmaf(pred_dt, acc[3u * i + 2u], vel[3u * i + 2u]), 0.0f);
}

// W(r, h) and dW/dr(r, h) of the cubic spline for 0 <= q = r/h
__device__ __forceinline__ float cubic_w(float q, float inv_h) {
	const float k = (8.0f / kPi) * inv_h * inv_h * inv_h;
	const float t = 1.0f - q;
	const float inner = __builtin_fmaf(6.0f * q * q, q - 1.0f, 1.0f);
	return q <= 0.5f ? k * inner : (q <= 1.0f ? k * 2.0f * t * t * t : 0.0f);
}
__device__ __forceinline__ float cubic_dw(float q, float inv_h) {
	const float k = (8.0f / kPi) * inv_h * inv_h * inv_h * inv_h;
	const float t = 1.0f - q;
	const float inner = q * __builtin_fmaf(18.0f, q, -12.0f);
	return q <= 0.5f ? k * inner : (q <= 1.0f ? k * -6.0f * t * t : 0.0f);
}

// Both passes split j into slices over the grid (slice = blockIdx % slices,
// as the fast PP kernel does) so that a few thousand particles still give
// every SIMD several wavefronts; the neighbour test is 8 VALU operations per
// pair and the kernel arithmetic runs only for lanes with a neighbour.
__global__ __launch_bounds__(kBlock) void k_rho_cubic(const float4* __restrict__ q0, const float4* __restrict__ q1,
                                                      unsigned n_pad, unsigned slices, unsigned slice_len,
                                                      float* __restrict__ part) {
	const unsigned b = blockIdx.x;
	const unsigned sl = b % slices;
	const unsigned i = (b / slices) * kBlock + threadIdx.x;
	const float4 a = q0[i];
	const float inv_hi = q1[i].y;
	float s = 0.0f;
	const unsigned j0 = sl * slice_len;
	const unsigned j1 = min(j0 + slice_len, n_pad);
	for (unsigned jt = j0; jt < j1; jt += kTj) {
#pragma unroll 8
		for (unsigned j = jt; j < jt + kTj; ++j) {
			const float4 b4 = q0[j];
			const float dx = a.x - b4.x, dy = a.y - b4.y, dz = a.z - b4.z;
			const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
			if (r2 < fmaxf(a.w, b4.w)) {
				const float4 c = q1[j];
				const float r = __builtin_sqrtf(r2);
				const float w = cubic_w(r * inv_hi, inv_hi) + cubic_w(r * c.y, c.y);
				s = __builtin_fmaf(0.5f * c.x, w, s);
			}
		}
	}
	part[(size_t)sl * n_pad + i] = s;
}

// rho = sum of the slice partials in fixed order; P = (gamma - 1) rho u;
// P / rho^2 into q1.w for the force pass
__global__ __launch_bounds__(kBlock) void k_rho_cubic_combine(const float* __restrict__ part, unsigned n_pad,
                                                              unsigned slices, unsigned N, float gm1,
                                                              float4* __restrict__ q1, float* __restrict__ rho,
                                                              float* __restrict__ prs) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	double s = 0.0;
	for (unsigned k = 0; k < slices; ++k) s += (double)part[(size_t)k * n_pad + i];
	const float r = (float)s;
	float4 c = q1[i];
	const float p = gm1 * r * c.z;
	c.w = r > 0.0f ? p / (r * r) : 0.0f;
	q1[i] = c;
	rho[i] = r;
	prs[i] = p;
}

__global__ __launch_bounds__(kBlock) void k_force_cubic(const float4* __restrict__ q0, const float4* __restrict__ q1,
                                                        const float4* __restrict__ q2, unsigned n_pad, unsigned slices,
                                                        unsigned slice_len, float4* __restrict__ part) {
	const unsigned b = blockIdx.x;
	const unsigned sl = b % slices;
	const unsigned i = (b / slices) * kBlock + threadIdx.x;
	const float4 a = q0[i];
	const float4 ca = q1[i];
	const float4 va = q2[i];
	float ax = 0.0f, ay = 0.0f, az = 0.0f, du = 0.0f;
	const unsigned j0 = sl * slice_len;
	const unsigned j1 = min(j0 + slice_len, n_pad);
	for (unsigned jt = j0; jt < j1; jt += kTj) {
#pragma unroll 8
		for (unsigned j = jt; j < jt + kTj; ++j) {
			const float4 b4 = q0[j];
			const float dx = a.x - b4.x, dy = a.y - b4.y, dz = a.z - b4.z;
			const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
			if (r2 < fmaxf(a.w, b4.w) && r2 > 0.0f) {
				const float4 cb = q1[j];
				const float4 vb = q2[j];
				const float inv_r = __builtin_amdgcn_rsqf(r2);
				const float r = r2 * inv_r;
				const float dw = cubic_dw(r * ca.y, ca.y) + cubic_dw(r * cb.y, cb.y);
				const float f = 0.5f * cb.x * (ca.w + cb.w) * dw * inv_r;   // grad_i W = dW/dr d / r
				ax = __builtin_fmaf(-f, dx, ax);
				ay = __builtin_fmaf(-f, dy, ay);
				az = __builtin_fmaf(-f, dz, az);
				const float vd =
				    __builtin_fmaf(va.z - vb.z, dz, __builtin_fmaf(va.y - vb.y, dy, (va.x - vb.x) * dx));
				du = __builtin_fmaf(0.5f * f, vd, du);
			}
		}
	}
	part[(size_t)sl * n_pad + i] = make_float4(ax, ay, az, du);
}

__global__ __launch_bounds__(kBlock) void k_force_cubic_combine(const float4* __restrict__ part, unsigned n_pad,
                                                                unsigned slices, unsigned N, int acc_phys,
                                                                const float4* __restrict__ q1,
                                                                float* __restrict__ acc, float* __restrict__ dudt) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	double ax = 0.0, ay = 0.0, az = 0.0, du = 0.0;
	for (unsigned k = 0; k < slices; ++k) {
		const float4 p = part[(size_t)k * n_pad + i];
		ax += (double)p.x;
		ay += (double)p.y;
		az += (double)p.z;
		du += (double)p.w;
	}
	const double sc = acc_phys ? 1.0 : (double)q1[i].x;
	acc[3u * i + 0u] += (float)(sc * ax);
	acc[3u * i + 1u] += (float)(sc * ay);
	acc[3u * i + 2u] += (float)(sc * az);
	dudt[i] = (float)du;
}

__global__ __launch_bounds__(kBlock) void k_u_kick(float* __restrict__ u, const float* __restrict__ dudt, unsigned N,
                                                   float half_dt) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < N) u[i] += half_dt * dudt[i];
}

inline dim3 grid(unsigned n) { return dim3((n + kBlock - 1u) / kBlock); }

}  // namespace

extern "C" {

size_t sfd_sph_tmp_floats(unsigned n_pad) { return 12u * (size_t)n_pad + 16u; }

int sfd_sph_ref(const float* pos, const float* mas, const float* rad, unsigned N, int kahan, float* rho, float* prs,
                float* acc, float* tmp, hipStream_t s) {
	if (N == 0) return 0;
	float* a3 = tmp;
	float* hh = tmp + N;
	float* den = tmp + 2u * (size_t)N;
	float* con = tmp + 3u * (size_t)N;
	hipLaunchKernelGGL(k_tables_ref, grid(N), dim3(kBlock), 0, s, rad, N, a3, hh, den);
	if (kahan) hipLaunchKernelGGL(k_rho_ref<true>, grid(N), dim3(kBlock), 0, s, pos, mas, a3, hh, N, rho);
	else hipLaunchKernelGGL(k_rho_ref<false>, grid(N), dim3(kBlock), 0, s, pos, mas, a3, hh, N, rho);
	hipLaunchKernelGGL(k_eos_ref, grid(N), dim3(kBlock), 0, s, rho, N, prs, con);
	if (kahan) hipLaunchKernelGGL(k_force_ref<true>, grid(N), dim3(kBlock), 0, s, pos, mas, hh, den, con, N, acc);
	else hipLaunchKernelGGL(k_force_ref<false>, grid(N), dim3(kBlock), 0, s, pos, mas, hh, den, con, N, acc);
	return sf_hip_check(hipGetLastError(), "solve_sph", "launch");
}

int sfd_sph_cubic_pack(const float* pos, const float* vel, const float* acc, const float* mas, const float* rad,
                       const float* u, const float* dudt, unsigned N, unsigned n_pad, float pred_dt, float* tmp,
                       hipStream_t s) {
	if (N == 0) return 0;
	float4* q0 = reinterpret_cast<float4*>(tmp);
	hipLaunchKernelGGL(k_pack_cubic, grid(n_pad), dim3(kBlock), 0, s, pos, vel, acc, mas, rad, u, dudt, N, n_pad,
	                   pred_dt, q0, q0 + n_pad, q0 + 2u * (size_t)n_pad);
	return sf_hip_check(hipGetLastError(), "solve_sph", "launch");
}

// part: slices x n_pad float4 of scratch
int sfd_sph_cubic(unsigned N, unsigned n_pad, float gamma, int acc_phys, float* rho, float* prs, float* acc,
                  float* dudt, float* tmp, float* part, unsigned slices, hipStream_t s) {
	if (N == 0) return 0;
	if (n_pad % kTj != 0 || n_pad % kBlock != 0 || slices == 0) return -1;
	float4* q0 = reinterpret_cast<float4*>(tmp);
	float4* q1 = q0 + n_pad;
	float4* q2 = q0 + 2u * (size_t)n_pad;
	const unsigned tiles = n_pad / kTj;
	if (slices > tiles) slices = tiles;
	const unsigned slice_len = (tiles + slices - 1u) / slices * kTj;
	slices = (n_pad + slice_len - 1u) / slice_len;
	const dim3 g((n_pad / kBlock) * slices);
	const int tok = sf_prof_begin("sph_rho_cubic", s);
	hipLaunchKernelGGL(k_rho_cubic, g, dim3(kBlock), 0, s, q0, q1, n_pad, slices, slice_len, part);
	hipLaunchKernelGGL(k_rho_cubic_combine, grid(N), dim3(kBlock), 0, s, part, n_pad, slices, N, gamma - 1.0f, q1,
	                   rho, prs);
	sf_prof_end(tok, s);
	const int tok2 = sf_prof_begin("sph_force_cubic", s);
	float4* p4 = reinterpret_cast<float4*>(part);
	hipLaunchKernelGGL(k_force_cubic, g, dim3(kBlock), 0, s, q0, q1, q2, n_pad, slices, slice_len, p4);
	hipLaunchKernelGGL(k_force_cubic_combine, grid(N), dim3(kBlock), 0, s, p4, n_pad, slices, N, acc_phys, q1, acc,
	                   dudt);
	sf_prof_end(tok2, s);
	return sf_hip_check(hipGetLastError(), "solve_sph", "launch");
}

int sfd_sph_u_kick(float* u, const float* dudt, unsigned N, float half_dt, hipStream_t s) {
	if (N == 0) return 0;
	hipLaunchKernelGGL(k_u_kick, grid(N), dim3(kBlock), 0, s, u, dudt, N, half_dt);
	return sf_hip_check(hipGetLastError(), "sim_step", "launch");
}

}  // extern "C"
