// lbvh.hip -- the SF_TREE_LBVH Barnes-Hut tree: build on gfx950 (the walk is
// in lbvh_walk.hip).
//
// Build (every step, O(N), bandwidth-bound):
//  1. FP64 bounding box (min/max are exact, so any reduction order gives the
//     same cube).
//  2. 63-bit Morton codes: q = floor((x - lo) * (2^21 / extent)) in FP64,
//     clamped to [0, 2^21), bits interleaved x-y-z from the top.  Exactly
//     reproducible by a CPU restatement.
//  3. Stable LSD radix sort of (code, index), 8 passes of 8-bit digits, with
//     per-tile histograms, one scan, and a ballot-based stable local rank.
//  4. Binary radix tree from the sorted codes (Karras 2012): internal node i
//     covers a contiguous range of the sorted particles; equal codes are
//     separated by their sorted position.
//  5. Monopole aggregation bottom-up with one integer arrival counter per
//     internal node; the second child to arrive combines left and right in a
//     fixed order, so the result does not depend on scheduling (no float
//     atomics).
//
// Walk records (k_records) carry, per node, the monopole, the squared
// critical radius for the current theta and either the particle range or the
// two children.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"
#include "tree_dev.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kTile = 4096;          // radix sort tile (256 threads x 16)
constexpr unsigned kRadixRounds = kTile / kBlock;

__global__ __launch_bounds__(kBlock) void k_fill_u32(unsigned* __restrict__ p, unsigned n, unsigned v) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, unsigned n) {
	if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------- bbox
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const float* __restrict__ pos, unsigned N,
                                                         double* __restrict__ part) {
	__shared__ double sh[6][kBlock];
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
		for (int a = 0; a < 3; ++a) {
			const double v = pos[3u * i + a];
			lo[a] = fmin(lo[a], v);
			hi[a] = fmax(hi[a], v);
		}
	}
	for (int a = 0; a < 3; ++a) {
		sh[a][threadIdx.x] = lo[a];
		sh[3 + a][threadIdx.x] = hi[a];
	}
	__syncthreads();
	for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
			for (int a = 0; a < 3; ++a) {
				sh[a][threadIdx.x] = fmin(sh[a][threadIdx.x], sh[a][threadIdx.x + h]);
				sh[3 + a][threadIdx.x] = fmax(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + h]);
			}
		}
		__syncthreads();
	}
	if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// box[0..2] = lo, box[3] = 2^21 / extent (extent = longest side, widened so
// the top cell is half-open)
__global__ void k_bbox_final(const double* __restrict__ part, unsigned nb, double* __restrict__ box) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (unsigned b = 0; b < nb; ++b)
		for (int a = 0; a < 3; ++a) {
			lo[a] = fmin(lo[a], part[b * 6 + a]);
			hi[a] = fmax(hi[a], part[b * 6 + 3 + a]);
		}
	double ext = 0.0;
	for (int a = 0; a < 3; ++a) ext = fmax(ext, hi[a] - lo[a]);
	if (!(ext > 0.0)) ext = 1.0;
	ext *= 1.0 + 1.0e-9;
	for (int a = 0; a < 3; ++a) box[a] = lo[a];
	box[3] = 2097152.0 / ext;
}

// ---------------------------------------------------------------- Morton
__device__ __forceinline__ unsigned long long spread21(unsigned long long v) {
	v &= 0x1FFFFFull;
	v = (v | (v << 32)) & 0x1F00000000FFFFull;
	v = (v | (v << 16)) & 0x1F0000FF0000FFull;
	v = (v | (v << 8)) & 0x100F00F00F00F00Full;
	v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
	v = (v | (v << 2)) & 0x1249249249249249ull;
	return v;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ pos, unsigned N, unsigned n_sort,
                                                   const double* __restrict__ box, unsigned long long* __restrict__ key,
                                                   unsigned* __restrict__ val) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_sort) return;
	if (i >= N) {
		key[i] = ~0ull;
		val[i] = i;
		return;
	}
	const double scale = box[3];
	unsigned long long q[3];
	for (int a = 0; a < 3; ++a) {
		double t = ((double)pos[3u * i + a] - box[a]) * scale;
		t = fmin(fmax(t, 0.0), 2097151.0);
		q[a] = (unsigned long long)t;
	}
	key[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
	val[i] = i;
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kBlock) void k_radix_hist(const unsigned long long* __restrict__ key, unsigned shift,
                                                       unsigned ntiles, unsigned* __restrict__ hist) {
	__shared__ unsigned h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	const unsigned base = blockIdx.x * kTile;
	for (unsigned r = 0; r < kRadixRounds; ++r) {
		const unsigned d = (unsigned)(key[base + r * kBlock + threadIdx.x] >> shift) & 255u;
		atomicAdd(&h[d], 1u);
	}
	__syncthreads();
	hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram (256 * ntiles entries) in one
// block: each thread scans a contiguous chunk, then the chunk totals.
__global__ __launch_bounds__(1024) void k_radix_scan(unsigned* __restrict__ hist, unsigned n) {
	__shared__ unsigned tot[1024];
	const unsigned chunk = (n + 1023) / 1024;
	const unsigned b = threadIdx.x * chunk, e = min(b + chunk, n);
	unsigned s = 0;
	for (unsigned k = b; k < e; ++k) s += hist[k];
	tot[threadIdx.x] = s;
	__syncthreads();
	for (unsigned off = 1; off < 1024; off <<= 1) {
		const unsigned v = threadIdx.x >= off ? tot[threadIdx.x - off] : 0u;
		__syncthreads();
		tot[threadIdx.x] += v;
		__syncthreads();
	}
	unsigned run = threadIdx.x ? tot[threadIdx.x - 1] : 0u;
	for (unsigned k = b; k < e; ++k) {
		const unsigned v = hist[k];
		hist[k] = run;
		run += v;
	}
}

// Stable scatter: tile elements are ranked in index order; within a round of
// 256 elements the rank among equal digits comes from a 64-lane match mask
// built with eight ballots.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const unsigned long long* __restrict__ key_in,
                                                          const unsigned* __restrict__ val_in, unsigned shift,
                                                          unsigned ntiles, const unsigned* __restrict__ offs,
                                                          unsigned long long* __restrict__ key_out,
                                                          unsigned* __restrict__ val_out) {
	__shared__ unsigned base_off[256];
	__shared__ unsigned run[256];
	__shared__ unsigned wcount[kBlock / 64][256];
	const unsigned t = threadIdx.x, lane = t & 63u, w = t >> 6;
	base_off[t] = offs[t * ntiles + blockIdx.x];
	run[t] = 0;
	for (unsigned q = 0; q < kBlock / 64; ++q) wcount[q][t] = 0;
	__syncthreads();
	const unsigned long long lt = (1ull << lane) - 1ull;
	for (unsigned r = 0; r < kRadixRounds; ++r) {
		const unsigned idx = blockIdx.x * kTile + r * kBlock + t;
		const unsigned long long k = key_in[idx];
		const unsigned v = val_in[idx];
		const unsigned d = (unsigned)(k >> shift) & 255u;
		unsigned long long match = ~0ull;
		for (int bit = 0; bit < 8; ++bit) {
			const unsigned long long b = __ballot((d >> bit) & 1u);
			match &= ((d >> bit) & 1u) ? b : ~b;
		}
		const unsigned rank = __popcll(match & lt);
		if (rank == 0) wcount[w][d] = __popcll(match);
		__syncthreads();
		unsigned before = run[d];
		for (unsigned q = 0; q < w; ++q) before += wcount[q][d];
		const unsigned dst = base_off[d] + before + rank;
		key_out[dst] = k;
		val_out[dst] = v;
		__syncthreads();
		unsigned add = 0;
		for (unsigned q = 0; q < kBlock / 64; ++q) {
			add += wcount[q][t];
			wcount[q][t] = 0;
		}
		run[t] += add;
		__syncthreads();
	}
	(void)ntiles;
}

// ---------------------------------------------------------------- tree
__global__ __launch_bounds__(kBlock) void k_gather_sorted(const float* __restrict__ pos, const float* __restrict__ mas,
                                                          const unsigned* __restrict__ order, unsigned N,
                                                          float4* __restrict__ body) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	const unsigned i = order[p];
	body[p] = make_float4(pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u], mas[i]);
}

__device__ __forceinline__ int delta(const unsigned long long* __restrict__ key, int N, int i, int j) {
	if (j < 0 || j >= N) return -1;
	const unsigned long long a = key[i], b = key[j];
	if (a != b) return __clzll((long long)(a ^ b));
	return 64 + __clz((int)((unsigned)i ^ (unsigned)j));
}

// Internal nodes 0..N-2; node id >= N-1 is the leaf of sorted particle
// id - (N-1).  node_lr holds (left, right), node_range (first, count).
__global__ __launch_bounds__(kBlock) void k_karras(const unsigned long long* __restrict__ key, int N,
                                                   int2* __restrict__ node_lr, int2* __restrict__ node_range,
                                                   int* __restrict__ parent) {
	const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (i >= N - 1) return;
	const int d = (delta(key, N, i, i + 1) - delta(key, N, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = delta(key, N, i, i - d);
	int lmax = 2;
	while (delta(key, N, i, i + lmax * d) > dmin) lmax <<= 1;
	int l = 0;
	for (int t = lmax >> 1; t >= 1; t >>= 1)
		if (delta(key, N, i, i + (l + t) * d) > dmin) l += t;
	const int j = i + l * d;
	const int dnode = delta(key, N, i, j);
	int s = 0;
	for (int div = 2;; div <<= 1) {
		const int t = (l + div - 1) / div;
		if (delta(key, N, i, i + (s + t) * d) > dnode) s += t;
		if (t <= 1) break;
	}
	const int gamma = i + s * d + min(d, 0);
	const int lo = min(i, j), hi = max(i, j);
	const int left = (lo == gamma) ? (N - 1) + gamma : gamma;
	const int right = (hi == gamma + 1) ? (N - 1) + gamma + 1 : gamma + 1;
	node_lr[i] = make_int2(left, right);
	node_range[i] = make_int2(lo, hi - lo + 1);
	parent[left] = i;
	parent[right] = i;
	if (i == 0) parent[0] = -1;
}

// node_cm: (com x, y, z, mass); node_box: (lo x, y, z, hi x) (hi y, hi z in node_box2)
__global__ __launch_bounds__(kBlock) void k_aggregate(const float4* __restrict__ body, int N,
                                                      const int2* __restrict__ node_lr, const int* __restrict__ parent,
                                                      unsigned* __restrict__ arrive, float4* __restrict__ node_cm,
                                                      float4* __restrict__ box_lo, float4* __restrict__ box_hi) {
	const int p = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (p >= N) return;
	const int leaf = (N - 1) + p;
	const float4 b = body[p];
	node_cm[leaf] = b;
	box_lo[leaf] = make_float4(b.x, b.y, b.z, 0.0f);
	box_hi[leaf] = make_float4(b.x, b.y, b.z, 0.0f);
	__threadfence();
	int n = parent[leaf];
	// a radix tree over 63-bit codes plus 32-bit positions is at most 96
	// levels deep; the step cap only guards against a malformed tree
	for (int step = 0; n >= 0 && n < N - 1 && step < 128; ++step) {
		const unsigned old = atomicAdd(&arrive[n], 1u);
		if (old != 1u) return;        // first arrival: the other child is not done yet
		__threadfence();
		const int2 lr = node_lr[n];
		// the agent-scope fences around the arrival counter make the sibling's
		// writes (possibly from another XCD) visible to these loads
		const float4 a = node_cm[lr.x];
		const float4 c = node_cm[lr.y];
		const float4 alo = box_lo[lr.x];
		const float4 clo = box_lo[lr.y];
		const float4 ahi = box_hi[lr.x];
		const float4 chi = box_hi[lr.y];
		const float m = a.w + c.w;
		float4 cm;
		if (m > 0.0f) {
			const float inv = 1.0f / m;
			cm = make_float4((a.w * a.x + c.w * c.x) * inv, (a.w * a.y + c.w * c.y) * inv,
			                 (a.w * a.z + c.w * c.z) * inv, m);
		} else {
			cm = make_float4(0.5f * (a.x + c.x), 0.5f * (a.y + c.y), 0.5f * (a.z + c.z), 0.0f);
		}
		node_cm[n] = cm;
		box_lo[n] = make_float4(fminf(alo.x, clo.x), fminf(alo.y, clo.y), fminf(alo.z, clo.z), 0.0f);
		box_hi[n] = make_float4(fmaxf(ahi.x, chi.x), fmaxf(ahi.y, chi.y), fmaxf(ahi.z, chi.z), 0.0f);
		__threadfence();
		n = parent[n];
	}
}

// Walk records: rec_a = (com, mass), rec_b = (r_crit^2, a, b, count) where
// (a, b) = (first, count) for nodes of at most leaf_size particles and
// (left, right) otherwise; leaves of one particle carry count 1.
__global__ __launch_bounds__(kBlock) void k_records(int N, const int2* __restrict__ node_lr,
                                                    const int2* __restrict__ node_range,
                                                    const float4* __restrict__ node_cm,
                                                    const float4* __restrict__ box_lo, const float4* __restrict__ box_hi,
                                                    float inv_theta, int leaf_size, float4* __restrict__ rec_a,
                                                    int4* __restrict__ rec_b) {
	const int n = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (n >= 2 * N - 1) return;
	const float4 cm = node_cm[n];
	rec_a[n] = cm;
	int first, count, l = -1, r = -1;
	if (n < N - 1) {
		const int2 rg = node_range[n];
		first = rg.x;
		count = rg.y;
		l = node_lr[n].x;
		r = node_lr[n].y;
	} else {
		first = n - (N - 1);
		count = 1;
	}
	const float4 lo = box_lo[n], hi = box_hi[n];
	const float sx = hi.x - lo.x, sy = hi.y - lo.y, sz = hi.z - lo.z;
	const float side = fmaxf(sx, fmaxf(sy, sz));
	const float ox = cm.x - 0.5f * (lo.x + hi.x), oy = cm.y - 0.5f * (lo.y + hi.y), oz = cm.z - 0.5f * (lo.z + hi.z);
	const float off = sqrtf(ox * ox + oy * oy + oz * oz);
	const float rc = side * inv_theta + off;
	const bool leafy = count <= leaf_size;
	// w carries the particle count, negated for nodes that are opened into
	// their children (whose ids are then in y, z)
	rec_b[n] = make_int4(__float_as_int(rc * rc), leafy ? first : l, leafy ? count : r, leafy ? count : -count);
}

}  // namespace

extern "C" {

unsigned sfd_lbvh_bbox_blocks(void) { return 256u; }

int sfd_lbvh_keys(const float* pos, unsigned N, unsigned n_sort, double* bbox_part, double* box,
                  unsigned long long* key, unsigned* val, hipStream_t s) {
	const unsigned nb = sfd_lbvh_bbox_blocks();
	if (N == 0 || n_sort % kTile != 0) return -1;
	hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kBlock), 0, s, pos, N, bbox_part);
	hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, bbox_part, nb, box);
	hipLaunchKernelGGL(k_morton, dim3((n_sort + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, N, n_sort, box, key,
	                   val);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_keys", "launch");
}

unsigned sfd_radix_tile(void) { return kTile; }

// Sorts n_sort (a multiple of the tile) pairs; the result ends in (key0,
// val0) after the eight passes.
int sfd_radix_sort(unsigned long long* key0, unsigned* val0, unsigned long long* key1, unsigned* val1,
                   unsigned n_sort, unsigned* hist, hipStream_t s) {
	const unsigned ntiles = n_sort / kTile;
	for (unsigned pass = 0; pass < 8; ++pass) {
		const unsigned shift = 8u * pass;
		unsigned long long* kin = (pass & 1) ? key1 : key0;
		unsigned* vin = (pass & 1) ? val1 : val0;
		unsigned long long* kout = (pass & 1) ? key0 : key1;
		unsigned* vout = (pass & 1) ? val0 : val1;
		hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, s, kin, shift, ntiles, hist);
		hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, hist, 256u * ntiles);
		hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kBlock), 0, s, kin, vin, shift, ntiles, hist, kout, vout);
	}
	return sf_hip_check(hipGetLastError(), "sfd_radix_sort", "launch");
}

int sfd_lbvh_topology(const float* pos, const float* mas, const unsigned long long* key, const unsigned* order,
                      unsigned N, void* body, void* node_lr, void* node_range, int* parent, unsigned* arrive,
                      void* node_cm, void* box_lo, void* box_hi, hipStream_t s) {
	const unsigned g = (N + kBlock - 1) / kBlock;
	float4* b4 = static_cast<float4*>(body);
	int2* lr = static_cast<int2*>(node_lr);
	int2* rg = static_cast<int2*>(node_range);
	const int tok = sf_prof_begin("lbvh_topology", s);
	hipLaunchKernelGGL(k_gather_sorted, dim3(g), dim3(kBlock), 0, s, pos, mas, order, N, b4);
	if (N > 1) {
		hipLaunchKernelGGL(k_karras, dim3((N - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, key, (int)N, lr, rg,
		                   parent);
		hipLaunchKernelGGL(k_fill_u32, dim3((N - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, arrive, N - 1, 0u);
	} else {
		hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, reinterpret_cast<unsigned*>(parent), 1u,
		                   0xFFFFFFFFu);
	}
	hipLaunchKernelGGL(k_aggregate, dim3(g), dim3(kBlock), 0, s, b4, (int)N, lr, parent, arrive,
	                   static_cast<float4*>(node_cm), static_cast<float4*>(box_lo), static_cast<float4*>(box_hi));
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_topology", "launch");
}

int sfd_lbvh_records(unsigned N, const void* node_lr, const void* node_range, const void* node_cm, const void* box_lo,
                     const void* box_hi, float theta, int leaf_size, void* rec_a, void* rec_b, hipStream_t s) {
	const unsigned nn = 2u * N - 1u;
	const float inv_theta = theta > 0.0f ? 1.0f / theta : 3.0e38f;
	hipLaunchKernelGGL(k_records, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (int)N,
	                   static_cast<const int2*>(node_lr), static_cast<const int2*>(node_range),
	                   static_cast<const float4*>(node_cm), static_cast<const float4*>(box_lo),
	                   static_cast<const float4*>(box_hi), inv_theta, leaf_size, static_cast<float4*>(rec_a),
	                   static_cast<int4*>(rec_b));
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_records", "launch");
}


}  // extern "C"
