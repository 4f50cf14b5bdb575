// sph_grid.hip -- pair search for the cubic-spline SPH model through a
// hashed cell grid (no reference counterpart: the reference tests every
// pair, src/sph.c:35-276).
//
// A pair (i, j) contributes when r^2 < max(h_i^2, h_j^2), so every partner
// of i lies within h_max of it.  Cells are cubes of side c slightly above
// h_max: partners sit in the 27 cells around i's own.  Cell coordinates
// come from floor(x / c) in FP64 (so rounding cannot move two particles
// closer than c more than one cell apart), clamped to 21 bits (clamping
// keeps neighbouring cells neighbouring), and hash into a power-of-two table
// of buckets.  Particles are sorted by bucket with the stable radix sort and
// their packed records gathered into that order; a bucket may hold several
// cells, so a candidate counts only when its cell code equals the target
// cell's.
//
// Each particle's 27 cells are split over 9 lanes (one (dz, dy) row of three
// cells each), which gives a 65536-particle system nine wavefronts per SIMD;
// the nine row partials are added in fixed row order, so results do not
// depend on scheduling.  The pair arithmetic is the all-pairs kernels'
// (device/sph.hip); only the order of each particle's sum differs.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"
#include "tree_dev.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned long long kNone = ~0ull;
constexpr unsigned kRows = 9;                 // (dz, dy) rows of three cells per particle
constexpr int kCellMax = (1 << 20) - 1;       // cell coordinates in [-2^20, 2^20 - 1]
constexpr float kPi = 3.14159265358979323846f;
constexpr unsigned kAlign = 256;

__device__ __forceinline__ int cell_coord(float x, double inv_c) {
	double t = floor((double)x * inv_c);
	t = fmin(fmax(t, -(double)(kCellMax + 1)), (double)kCellMax);   // NaN goes to the low end
	return (int)t;
}

__device__ __forceinline__ unsigned cell_hash(int cx, int cy, int cz, unsigned mask) {
	return (((unsigned)cx * 73856093u) ^ ((unsigned)cy * 19349663u) ^ ((unsigned)cz * 83492791u)) & mask;
}

__device__ __forceinline__ unsigned long long cell_code(int cx, int cy, int cz) {
	return ((unsigned long long)((unsigned)cx & 0x1FFFFFu) << 42) |
	       ((unsigned long long)((unsigned)cy & 0x1FFFFFu) << 21) | (unsigned long long)((unsigned)cz & 0x1FFFFFu);
}

__device__ __forceinline__ int sext21(unsigned long long v) {
	return (int)((unsigned)(v & 0x1FFFFFu) << 11) >> 11;
}

// W(r, h) and dW/dr(r, h) of the cubic spline, as in device/sph.hip
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

// hm[0] = bits of max |h| (non-negative floats order as integers)
__global__ __launch_bounds__(kBlock) void k_hmax(const float* __restrict__ rad, unsigned N, unsigned* __restrict__ hm) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	float h = i < N ? fabsf(rad[i]) : 0.0f;
	if (!(h <= 3.0e38f)) h = 0.0f;   // NaN and inf take no part in the search radius
	unsigned v = __float_as_uint(h);
	for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o));
	if ((threadIdx.x & 63u) == 0u) atomicMax(hm, v);
}

// 1 / c with c = h_max (1 + 2^-10), the margin covering the FP32 rounding of
// the pair test; unit cells when every h is zero
__global__ void k_cell_scale(const unsigned* __restrict__ hm, double* __restrict__ inv_c) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	const double h = (double)__uint_as_float(hm[0]);
	inv_c[0] = h > 0.0 ? 1.0 / (h * (1.0 + 1.0 / 1024.0)) : 1.0;
}

__global__ void k_zero_u32(unsigned* __restrict__ x) {
	if (blockIdx.x == 0 && threadIdx.x == 0) x[0] = 0u;
}

// key = bucket of the particle's cell, val = index; padding keys kNone
__global__ __launch_bounds__(kBlock) void k_grid_keys(const float4* __restrict__ q0, unsigned N, unsigned n_sort,
                                                      const double* __restrict__ inv_c, unsigned mask,
                                                      unsigned long long* __restrict__ key,
                                                      unsigned* __restrict__ val) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_sort) return;
	unsigned long long k = kNone;
	if (i < N) {
		const float4 a = q0[i];
		const double ic = inv_c[0];
		k = cell_hash(cell_coord(a.x, ic), cell_coord(a.y, ic), cell_coord(a.z, ic), mask);
	}
	key[i] = k;
	val[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_bucket_clear(uint2* __restrict__ bucket, unsigned T) {
	const unsigned b = blockIdx.x * kBlock + threadIdx.x;
	if (b < T) bucket[b] = make_uint2(0u, 0u);
}

// [first, end) of each bucket's run of sorted keys
__global__ __launch_bounds__(kBlock) void k_bucket_mark(const unsigned long long* __restrict__ key, unsigned N,
                                                        uint2* __restrict__ bucket) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	const unsigned long long k = key[p];
	const unsigned b = (unsigned)k;
	if (p == 0u || key[p - 1u] != k) bucket[b].x = p;
	if (p + 1u == N || key[p + 1u] != k) bucket[b].y = p + 1u;
}

// packed records and cell codes in bucket order
__global__ __launch_bounds__(kBlock) void k_grid_gather(const float4* __restrict__ q0, const float4* __restrict__ q1,
                                                        const float4* __restrict__ q2, const unsigned* __restrict__ val,
                                                        unsigned N, const double* __restrict__ inv_c,
                                                        float4* __restrict__ s0, float4* __restrict__ s1,
                                                        float4* __restrict__ s2,
                                                        unsigned long long* __restrict__ scode) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	const unsigned i = val[p];
	const float4 a = q0[i];
	s0[p] = a;
	s1[p] = q1[i];
	s2[p] = q2[i];
	const double ic = inv_c[0];
	scode[p] = cell_code(cell_coord(a.x, ic), cell_coord(a.y, ic), cell_coord(a.z, ic));
}

// Lane (row, p): the three cells (cx - 1 .. cx + 1, cy + dy, cz + dz) of
// sorted particle p; calls f(j) for every candidate j in those cells.
template <class F>
__device__ __forceinline__ void for_row(unsigned row, unsigned long long own, const uint2* __restrict__ bucket,
                                        const unsigned long long* __restrict__ scode, unsigned mask, F&& f) {
	const int cx = sext21(own >> 42), cy = sext21(own >> 21) + (int)(row % 3u) - 1,
	          cz = sext21(own) + (int)(row / 3u) - 1;
	if (cy < -(kCellMax + 1) || cy > kCellMax || cz < -(kCellMax + 1) || cz > kCellMax) return;
	for (int dx = -1; dx <= 1; ++dx) {
		const int x = cx + dx;
		if (x < -(kCellMax + 1) || x > kCellMax) continue;
		const unsigned long long target = cell_code(x, cy, cz);
		const uint2 r = bucket[cell_hash(x, cy, cz, mask)];
		for (unsigned j = r.x; j < r.y; ++j)
			if (scode[j] == target) f(j);
	}
}

__global__ __launch_bounds__(kBlock) void k_rho_grid(const float4* __restrict__ s0, const float4* __restrict__ s1,
                                                     const unsigned long long* __restrict__ scode,
                                                     const uint2* __restrict__ bucket, unsigned N, unsigned mask,
                                                     float* __restrict__ part) {
	const unsigned g = blockIdx.x * kBlock + threadIdx.x;
	if (g >= kRows * N) return;
	const unsigned row = g / N, p = g - row * N;
	const float4 a = s0[p];
	const float inv_hi = s1[p].y;
	float s = 0.0f;
	for_row(row, scode[p], bucket, scode, mask, [&](unsigned j) {
		const float4 b4 = s0[j];
		const float dx = a.x - b4.x, dy = a.y - b4.y, dz = a.z - b4.z;
		const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
		if (r2 < fmaxf(a.w, b4.w)) {
			const float4 c = s1[j];
			const float r = __builtin_sqrtf(r2);
			const float w = cubic_w(r * inv_hi, inv_hi) + cubic_w(r * c.y, c.y);
			s = __builtin_fmaf(0.5f * c.x, w, s);
		}
	});
	part[(size_t)row * N + p] = s;
}

// rho = the row partials in row order; P = (gamma - 1) rho u; P / rho^2 into
// the sorted q1.w for the force pass
__global__ __launch_bounds__(kBlock) void k_rho_grid_finish(const float* __restrict__ part,
                                                            const unsigned* __restrict__ val, unsigned N, float gm1,
                                                            float4* __restrict__ s1, float* __restrict__ rho,
                                                            float* __restrict__ prs) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	float r = 0.0f;
	for (unsigned k = 0; k < kRows; ++k) r += part[(size_t)k * N + p];
	float4 c = s1[p];
	const float pr = gm1 * r * c.z;
	c.w = r > 0.0f ? pr / (r * r) : 0.0f;
	s1[p] = c;
	const unsigned i = val[p];
	rho[i] = r;
	prs[i] = pr;
}

__global__ __launch_bounds__(kBlock) void k_force_grid(const float4* __restrict__ s0, const float4* __restrict__ s1,
                                                       const float4* __restrict__ s2,
                                                       const unsigned long long* __restrict__ scode,
                                                       const uint2* __restrict__ bucket, unsigned N, unsigned mask,
                                                       float4* __restrict__ part) {
	const unsigned g = blockIdx.x * kBlock + threadIdx.x;
	if (g >= kRows * N) return;
	const unsigned row = g / N, p = g - row * N;
	const float4 a = s0[p];
	const float4 ca = s1[p];
	const float4 va = s2[p];
	float ax = 0.0f, ay = 0.0f, az = 0.0f, du = 0.0f;
	for_row(row, scode[p], bucket, scode, mask, [&](unsigned j) {
		const float4 b4 = s0[j];
		const float dx = a.x - b4.x, dy = a.y - b4.y, dz = a.z - b4.z;
		const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
		if (r2 < fmaxf(a.w, b4.w) && r2 > 0.0f) {
			const float4 cb = s1[j];
			const float4 vb = s2[j];
			const float inv_r = __builtin_amdgcn_rsqf(r2);
			const float r = r2 * inv_r;
			const float dw = cubic_dw(r * ca.y, ca.y) + cubic_dw(r * cb.y, cb.y);
			const float f = 0.5f * cb.x * (ca.w + cb.w) * dw * inv_r;   // grad_i W = dW/dr d / r
			ax = __builtin_fmaf(-f, dx, ax);
			ay = __builtin_fmaf(-f, dy, ay);
			az = __builtin_fmaf(-f, dz, az);
			const float vd = __builtin_fmaf(va.z - vb.z, dz, __builtin_fmaf(va.y - vb.y, dy, (va.x - vb.x) * dx));
			du = __builtin_fmaf(0.5f * f, vd, du);
		}
	});
	part[(size_t)row * N + p] = make_float4(ax, ay, az, du);
}

// acc_i += m-scaled sum of the row partials (row order), dudt_i = their du
__global__ __launch_bounds__(kBlock) void k_force_grid_finish(const float4* __restrict__ part,
                                                              const unsigned* __restrict__ val,
                                                              const float4* __restrict__ s1, unsigned N, int acc_phys,
                                                              float* __restrict__ acc, float* __restrict__ dudt) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	float ax = 0.0f, ay = 0.0f, az = 0.0f, du = 0.0f;
	for (unsigned k = 0; k < kRows; ++k) {
		const float4 q = part[(size_t)k * N + p];
		ax += q.x;
		ay += q.y;
		az += q.z;
		du += q.w;
	}
	const float sc = acc_phys ? 1.0f : s1[p].x;
	const unsigned i = val[p];
	acc[3u * i + 0u] += sc * ax;
	acc[3u * i + 1u] += sc * ay;
	acc[3u * i + 2u] += sc * az;
	dudt[i] = du;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + kBlock - 1u) / kBlock); }
inline size_t align_up(size_t n) { return (n + kAlign - 1u) / kAlign * kAlign; }

unsigned table_size(unsigned N) {
	unsigned T = 1024u;
	while (T < 2u * N && T < (1u << 24)) T <<= 1;
	return T;
}

enum Plane { KEY0, KEY1, VAL0, VAL1, HIST, BUCKET, S0, S1, S2, SCODE, PART, HM, INVC, kPlanes };

struct Layout {
	unsigned n_sort, T;
	size_t off[kPlanes];
	size_t bytes;
};

Layout layout(unsigned N) {
	Layout L;
	const unsigned tile = sfd_radix_tile();
	L.n_sort = (N + tile - 1u) / tile * tile;
	if (L.n_sort == 0) L.n_sort = tile;
	L.T = table_size(N);
	size_t sz[kPlanes];
	sz[KEY0] = sz[KEY1] = 8u * (size_t)L.n_sort;
	sz[VAL0] = sz[VAL1] = 4u * (size_t)L.n_sort;
	sz[HIST] = 4u * (size_t)sfd_radix_hist_entries(L.n_sort);
	sz[BUCKET] = 8u * (size_t)L.T;
	sz[S0] = sz[S1] = sz[S2] = 16u * (size_t)N;
	sz[SCODE] = 8u * (size_t)N;
	sz[PART] = 16u * (size_t)kRows * N;
	sz[HM] = 4u;
	sz[INVC] = 8u;
	size_t o = 0;
	for (int k = 0; k < kPlanes; ++k) {
		L.off[k] = o;
		o += align_up(sz[k] ? sz[k] : 1u);
	}
	L.bytes = o;
	return L;
}

template <class T>
T* plane(void* scratch, const Layout& L, Plane k) {
	return reinterpret_cast<T*>(static_cast<char*>(scratch) + L.off[k]);
}

}  // namespace

extern "C" {

size_t sfd_sph_grid_bytes(unsigned N) { return layout(N).bytes; }

// rho, prs, acc (+=) and dudt of the cubic model from the packed records
// (tmp: q0, q1, q2 of n_pad float4 each, as sfd_sph_cubic_pack writes them)
// through the cell grid; scratch: sfd_sph_grid_bytes(N) bytes.
int sfd_sph_cubic_grid(const float* rad, unsigned N, unsigned n_pad, float gamma, int acc_phys, float* rho, float* prs,
                       float* acc, float* dudt, float* tmp, void* scratch, hipStream_t s) {
	if (N == 0) return 0;
	const Layout L = layout(N);
	float4* q0 = reinterpret_cast<float4*>(tmp);
	float4* q1 = q0 + n_pad;
	float4* q2 = q0 + 2u * (size_t)n_pad;
	unsigned long long* key0 = plane<unsigned long long>(scratch, L, KEY0);
	unsigned long long* key1 = plane<unsigned long long>(scratch, L, KEY1);
	unsigned* val0 = plane<unsigned>(scratch, L, VAL0);
	unsigned* val1 = plane<unsigned>(scratch, L, VAL1);
	unsigned* hist = plane<unsigned>(scratch, L, HIST);
	uint2* bucket = plane<uint2>(scratch, L, BUCKET);
	float4* s0 = plane<float4>(scratch, L, S0);
	float4* s1 = plane<float4>(scratch, L, S1);
	float4* s2 = plane<float4>(scratch, L, S2);
	unsigned long long* scode = plane<unsigned long long>(scratch, L, SCODE);
	float* part = plane<float>(scratch, L, PART);
	unsigned* hm = plane<unsigned>(scratch, L, HM);
	double* inv_c = plane<double>(scratch, L, INVC);
	const unsigned mask = L.T - 1u;
	// passes for bucket keys below T, kept below the padding key's low bytes
	unsigned passes = 1u;
	while (passes < 8u && (unsigned long long)mask >= (1ull << (8u * passes)) - 1ull) ++passes;

	const int tok = sf_prof_begin("sph_grid_sort", s);
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, hm);
	hipLaunchKernelGGL(k_hmax, dim3(blocks(N)), dim3(kBlock), 0, s, rad, N, hm);
	hipLaunchKernelGGL(k_cell_scale, dim3(1), dim3(64), 0, s, hm, inv_c);
	hipLaunchKernelGGL(k_grid_keys, dim3(blocks(L.n_sort)), dim3(kBlock), 0, s, q0, N, L.n_sort, inv_c, mask, key0,
	                   val0);
	if (sfd_radix_sort_passes(key0, val0, key1, val1, L.n_sort, hist, passes, s)) return -1;
	unsigned long long* key = (passes & 1u) ? key1 : key0;
	unsigned* val = (passes & 1u) ? val1 : val0;
	hipLaunchKernelGGL(k_bucket_clear, dim3(blocks(L.T)), dim3(kBlock), 0, s, bucket, L.T);
	hipLaunchKernelGGL(k_bucket_mark, dim3(blocks(N)), dim3(kBlock), 0, s, key, N, bucket);
	hipLaunchKernelGGL(k_grid_gather, dim3(blocks(N)), dim3(kBlock), 0, s, q0, q1, q2, val, N, inv_c, s0, s1, s2,
	                   scode);
	sf_prof_end(tok, s);
	const int tok1 = sf_prof_begin("sph_rho_grid", s);
	hipLaunchKernelGGL(k_rho_grid, dim3(blocks((size_t)kRows * N)), dim3(kBlock), 0, s, s0, s1, scode, bucket, N,
	                   mask, part);
	hipLaunchKernelGGL(k_rho_grid_finish, dim3(blocks(N)), dim3(kBlock), 0, s, part, val, N, gamma - 1.0f, s1, rho,
	                   prs);
	sf_prof_end(tok1, s);
	const int tok2 = sf_prof_begin("sph_force_grid", s);
	float4* p4 = reinterpret_cast<float4*>(part);
	hipLaunchKernelGGL(k_force_grid, dim3(blocks((size_t)kRows * N)), dim3(kBlock), 0, s, s0, s1, s2, scode, bucket,
	                   N, mask, p4);
	hipLaunchKernelGGL(k_force_grid_finish, dim3(blocks(N)), dim3(kBlock), 0, s, p4, val, s1, N, acc_phys, acc, dudt);
	sf_prof_end(tok2, s);
	return sf_hip_check(hipGetLastError(), "solve_sph", "grid launch");
}

}  // extern "C"
