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
//  5. Monopoles from FP64 prefix sums of (m, m x, m y, m z) over the sorted
//     particles: a node's mass and moments are differences of two prefix
//     entries, so no bottom-up pass, atomics or cross-XCD fences are needed,
//     and the fixed-shape scan makes the result deterministic.  A node's size
//     in the opening test is its particles' bounding box when it holds at
//     most 64 particles, else its Morton cell: the cell fixed by the common
//     prefix of its first and last codes.
//
// Walk records (k_records) carry, per node, the squared critical radius for
// the current theta, either the particle range or the two children, and the
// size term of the relative opening test.
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
__global__ __launch_bounds__(kBlock) void k_bbox_final(const double* __restrict__ part, unsigned nb,
                                                       double* __restrict__ box) {
	__shared__ double sh[6][kBlock];
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (unsigned b = threadIdx.x; b < nb; b += kBlock)
		for (int a = 0; a < 3; ++a) {
			lo[a] = fmin(lo[a], part[b * 6 + a]);
			hi[a] = fmax(hi[a], part[b * 6 + 3 + a]);
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
	if (threadIdx.x != 0) return;
	double ext = 0.0;
	for (int a = 0; a < 3; ++a) ext = fmax(ext, sh[3 + a][0] - sh[a][0]);
	if (!(ext > 0.0)) ext = 1.0;
	ext *= 1.0 + 1.0e-9;
	for (int a = 0; a < 3; ++a) box[a] = sh[a][0];
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

// Keys, and (x, y, z, m) packed in index order, so that the gather after the
// sort reads one 16-byte record per particle instead of two cache lines.
__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ pos, const float* __restrict__ mas,
                                                   unsigned N, unsigned n_sort, const double* __restrict__ box,
                                                   unsigned long long* __restrict__ key, unsigned* __restrict__ val,
                                                   float4* __restrict__ packed) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_sort) return;
	if (i >= N) {
		key[i] = ~0ull;
		val[i] = i;
		return;
	}
	const double scale = box[3];
	const float x = pos[3u * i], y = pos[3u * i + 1u], z = pos[3u * i + 2u];
	const float c[3] = {x, y, z};
	unsigned long long q[3];
	for (int a = 0; a < 3; ++a) {
		double t = ((double)c[a] - box[a]) * scale;
		t = fmin(fmax(t, 0.0), 2097151.0);
		q[a] = (unsigned long long)t;
	}
	key[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
	val[i] = i;
	packed[i] = make_float4(x, y, z, mas[i]);
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

// Exclusive scan of the digit-major histogram (256 * ntiles entries) in three
// passes: per-chunk totals (k_scan_u32_sums), an exclusive scan of the chunk
// totals in one block (k_radix_scan), and the scan of each chunk from its
// offset (k_scan_u32_down).  A chunk is 4096 entries, 16 per thread.
constexpr unsigned kScanU32Per = 16;
constexpr unsigned kScanU32Chunk = kBlock * kScanU32Per;

__global__ __launch_bounds__(kBlock) void k_scan_u32_sums(const unsigned* __restrict__ data, unsigned n,
                                                          unsigned* __restrict__ sums) {
	__shared__ unsigned ws[kBlock / 64];
	const unsigned base = blockIdx.x * kScanU32Chunk + threadIdx.x * kScanU32Per;
	unsigned s = 0;
#pragma unroll
	for (unsigned k = 0; k < kScanU32Per; ++k) s += base + k < n ? data[base + k] : 0u;
	for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
	if ((threadIdx.x & 63u) == 0) ws[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned t = 0;
		for (unsigned q = 0; q < kBlock / 64; ++q) t += ws[q];
		sums[blockIdx.x] = t;
	}
}

__global__ __launch_bounds__(kBlock) void k_scan_u32_down(unsigned* __restrict__ data, unsigned n,
                                                          const unsigned* __restrict__ offs) {
	__shared__ unsigned ws[kBlock / 64];
	const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	const unsigned base = blockIdx.x * kScanU32Chunk + threadIdx.x * kScanU32Per;
	unsigned v[kScanU32Per];
	unsigned s = 0;
#pragma unroll
	for (unsigned k = 0; k < kScanU32Per; ++k) {
		v[k] = base + k < n ? data[base + k] : 0u;
		s += v[k];
	}
	unsigned inc = s;
	for (int o = 1; o < 64; o <<= 1) {
		const unsigned t = __shfl_up(inc, o);
		if ((int)lane >= o) inc += t;
	}
	if (lane == 63u) ws[w] = inc;
	__syncthreads();
	unsigned run = offs[blockIdx.x] + inc - s;
	for (unsigned q = 0; q < w; ++q) run += ws[q];
#pragma unroll
	for (unsigned k = 0; k < kScanU32Per; ++k) {
		if (base + k < n) data[base + k] = run;
		run += v[k];
	}
}

// Exclusive scan of n entries in one block: each thread scans a contiguous
// chunk, then the chunk totals (the top level of the histogram scan).
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

// Stable scatter of one 4096-element tile.  Each wavefront ranks its 1024
// consecutive elements in index order (a 64-lane match mask from eight
// ballots per round, and per-wave running digit counts in LDS), the tile is
// staged in LDS in sorted order, and it is written out by consecutive
// threads, so that each digit's run goes to consecutive global addresses.
// The tile's digit counts are differences of consecutive entries of the
// scanned histogram (offs[256 * ntiles] would be n_total).
constexpr unsigned kScatterPer = kTile / kBlock;          // elements per thread
constexpr unsigned kWaveSpan = kTile / (kBlock / 64);     // elements per wavefront

__global__ __launch_bounds__(kBlock) void k_radix_scatter(const unsigned long long* __restrict__ key_in,
                                                          const unsigned* __restrict__ val_in, unsigned shift,
                                                          unsigned ntiles, const unsigned* __restrict__ offs,
                                                          unsigned n_total, unsigned long long* __restrict__ key_out,
                                                          unsigned* __restrict__ val_out) {
	__shared__ unsigned long long s_key[kTile];
	__shared__ unsigned s_val[kTile];
	__shared__ unsigned s_wcnt[kBlock / 64][256];   // running counts, then each wave's prefix
	__shared__ unsigned s_start[256];               // the digit's first slot in the tile
	__shared__ unsigned s_base[256];                // the digit's first global position of this tile
	__shared__ unsigned s_wsum[kBlock / 64];
	const unsigned t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const unsigned at = t * ntiles + blockIdx.x;
	const unsigned o = offs[at];
	const unsigned cnt = (at + 1u < 256u * ntiles ? offs[at + 1u] : n_total) - o;
	s_base[t] = o;
	// exclusive scan of the tile's 256 digit counts
	unsigned inc = cnt;
	for (int k = 1; k < 64; k <<= 1) {
		const unsigned v = __shfl_up(inc, (unsigned)k);
		if ((int)lane >= k) inc += v;
	}
	if (lane == 63u) s_wsum[w] = inc;
	for (unsigned q = 0; q < kBlock / 64; ++q) s_wcnt[q][t] = 0;
	__syncthreads();
	unsigned excl = inc - cnt;
	for (unsigned q = 0; q < w; ++q) excl += s_wsum[q];
	s_start[t] = excl;
	// rank: element r of this thread is tile element w * kWaveSpan + 64 r + lane
	const unsigned long long lt = (1ull << lane) - 1ull;
	const unsigned g0 = blockIdx.x * kTile + w * kWaveSpan + lane;
	unsigned long long kr[kScatterPer];
	unsigned vr[kScatterPer], pk[kScatterPer];
#pragma unroll
	for (unsigned r = 0; r < kScatterPer; ++r) {
		const unsigned long long k = key_in[g0 + 64u * r];
		const unsigned d = (unsigned)(k >> shift) & 255u;
		unsigned long long match = ~0ull;
		for (int bit = 0; bit < 8; ++bit) {
			const unsigned long long b = __ballot((d >> bit) & 1u);
			match &= ((d >> bit) & 1u) ? b : ~b;
		}
		const unsigned before = s_wcnt[w][d];
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		const unsigned rank = __popcll(match & lt);
		if (rank == 0u) s_wcnt[w][d] = before + (unsigned)__popcll(match);
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		kr[r] = k;
		vr[r] = val_in[g0 + 64u * r];
		pk[r] = (d << 16) | (before + rank);   // before + rank < kWaveSpan
	}
	__syncthreads();
	// each wave's offset within every digit
	unsigned run = 0;
	for (unsigned q = 0; q < kBlock / 64; ++q) {
		const unsigned c = s_wcnt[q][t];
		s_wcnt[q][t] = run;
		run += c;
	}
	__syncthreads();
#pragma unroll
	for (unsigned r = 0; r < kScatterPer; ++r) {
		const unsigned d = pk[r] >> 16;
		const unsigned slot = s_start[d] + s_wcnt[w][d] + (pk[r] & 0xFFFFu);
		s_key[slot] = kr[r];
		s_val[slot] = vr[r];
	}
	__syncthreads();
#pragma unroll
	for (unsigned r = 0; r < kScatterPer; ++r) {
		const unsigned slot = r * kBlock + t;
		const unsigned long long k = s_key[slot];
		const unsigned d = (unsigned)(k >> shift) & 255u;
		const unsigned dst = s_base[d] + slot - s_start[d];
		key_out[dst] = k;
		val_out[dst] = s_val[slot];
	}
}

// ---------------------------------------------------------------- tree
__global__ __launch_bounds__(kBlock) void k_gather_sorted(const float4* __restrict__ packed,
                                                          const unsigned* __restrict__ order, unsigned N,
                                                          float4* __restrict__ body) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (p >= N) return;
	body[p] = packed[order[p]];
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
                                                   int2* __restrict__ node_lr, int2* __restrict__ node_range) {
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
}

// ---- FP64 inclusive prefix sums over the sorted particles of (m, m x, m y,
// m z) and, for quadrupoles, of m x' x'^T (x' = x - ref, six components).
// Fixed shape: 16 consecutive elements per thread, a Hillis-Steele scan of
// the 256 thread sums per block, one block scanning the block sums.
struct d4 {
	double m, x, y, z;
	__device__ static d4 zero() { return {0.0, 0.0, 0.0, 0.0}; }
	__device__ static d4 of(float4 b, const double*) {
		const double m = b.w;
		return {m, m * (double)b.x, m * (double)b.y, m * (double)b.z};
	}
	__device__ d4 operator+(const d4& o) const { return {m + o.m, x + o.x, y + o.y, z + o.z}; }
	__device__ d4 shfl_up(unsigned o) const {
		return {__shfl_up(m, o), __shfl_up(x, o), __shfl_up(y, o), __shfl_up(z, o)};
	}
	__device__ d4 shfl(int src) const { return {__shfl(m, src), __shfl(x, src), __shfl(y, src), __shfl(z, src)}; }
};
struct d6 {   // xx, xy, xz, yy, yz, zz
	double v[6];
	__device__ static d6 zero() { return {{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}}; }
	__device__ static d6 of(float4 b, const double* ref) {
		const double m = b.w, x = (double)b.x - ref[0], y = (double)b.y - ref[1], z = (double)b.z - ref[2];
		return {{m * x * x, m * x * y, m * x * z, m * y * y, m * y * z, m * z * z}};
	}
	__device__ d6 operator+(const d6& o) const {
		d6 r;
		for (int k = 0; k < 6; ++k) r.v[k] = v[k] + o.v[k];
		return r;
	}
	__device__ d6 shfl_up(unsigned o) const {
		d6 r;
		for (int k = 0; k < 6; ++k) r.v[k] = __shfl_up(v[k], o);
		return r;
	}
	__device__ d6 shfl(int src) const {
		d6 r;
		for (int k = 0; k < 6; ++k) r.v[k] = __shfl(v[k], src);
		return r;
	}
};

constexpr unsigned kScanPer = 16;
constexpr unsigned kScanTile = kBlock * kScanPer;
constexpr unsigned kScanWaveSpan = kScanTile / (kBlock / 64);   // consecutive elements per wavefront

template <class R>
__device__ R block_exclusive_scan(R v, R* sh) {
	const unsigned t = threadIdx.x;
	sh[t] = v;
	__syncthreads();
	for (unsigned off = 1; off < kBlock; off <<= 1) {
		const R o = t >= off ? sh[t - off] : R::zero();
		__syncthreads();
		sh[t] = sh[t] + o;
		__syncthreads();
	}
	const R excl = t ? sh[t - 1] : R::zero();
	__syncthreads();
	return excl;
}

// Sum of a wavefront's kScanWaveSpan consecutive elements (read 64
// consecutive at a time), reduced across lanes in a fixed order; every lane
// returns it.
template <class R>
__device__ R wave_span_sum(const float4* __restrict__ body, unsigned N, const double* __restrict__ ref,
                           unsigned first) {
	const unsigned lane = threadIdx.x & 63u;
	R s = R::zero();
	for (unsigned r = 0; r < kScanPer; ++r) {
		const unsigned i = first + 64u * r + lane;
		if (i < N) s = s + R::of(body[i], ref);
	}
	for (unsigned o = 1; o < 64; o <<= 1) {
		const R t = s.shfl_up(o);
		if (lane >= o) s = s + t;
	}
	return s.shfl(63);
}

// Tile sums: each wavefront sums its kScanWaveSpan elements, then wave 0 adds
// the four in order.
template <class R>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const float4* __restrict__ body, unsigned N,
                                                       const double* __restrict__ ref, R* __restrict__ tile_sum) {
	__shared__ R sh[kBlock / 64];
	const unsigned w = threadIdx.x >> 6;
	const R s = wave_span_sum<R>(body, N, ref, blockIdx.x * kScanTile + w * kScanWaveSpan);
	if ((threadIdx.x & 63u) == 0u) sh[w] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		R t = sh[0];
		for (unsigned q = 1; q < kBlock / 64; ++q) t = t + sh[q];
		tile_sum[blockIdx.x] = t;
	}
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(R* __restrict__ tile_sum, unsigned ntiles) {
	__shared__ R sh[kBlock];
	const unsigned chunk = (ntiles + kBlock - 1) / kBlock;
	const unsigned b = threadIdx.x * chunk, e = min(b + chunk, ntiles);
	R s = R::zero();
	for (unsigned k = b; k < e; ++k) s = s + tile_sum[k];
	R run = block_exclusive_scan(s, sh);
	for (unsigned k = b; k < e; ++k) {
		const R v = tile_sum[k];
		tile_sum[k] = run;       // exclusive offset of tile k
		run = run + v;
	}
}

// Inclusive prefix of a tile: each wavefront starts from the tile offset plus
// the sums of the wavefronts before it, and scans its elements 64 at a time
// (a lane scan, then the round's total carried to the next round), so that
// reads and writes are of consecutive records.
template <class R>
__global__ __launch_bounds__(kBlock) void k_scan_write(const float4* __restrict__ body, unsigned N,
                                                       const double* __restrict__ ref, const R* __restrict__ tile_off,
                                                       R* __restrict__ prefix) {
	__shared__ R sh[kBlock / 64];
	const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	const unsigned first = blockIdx.x * kScanTile + w * kScanWaveSpan;
	const R total = wave_span_sum<R>(body, N, ref, first);
	if (lane == 0u) sh[w] = total;
	__syncthreads();
	R run = tile_off[blockIdx.x];
	for (unsigned q = 0; q < w; ++q) run = run + sh[q];
	for (unsigned r = 0; r < kScanPer; ++r) {
		const unsigned i = first + 64u * r + lane;
		R v = i < N ? R::of(body[i], ref) : R::zero();
		for (unsigned o = 1; o < 64; o <<= 1) {
			const R t = v.shfl_up(o);
			if (lane >= o) v = v + t;
		}
		if (i < N) prefix[i] = run + v;   // inclusive
		run = run + v.shfl(63);
	}
}

// Reference point of the second-moment prefix sums: the global centre of
// mass (the last first-moment prefix), so that the sums stay of the order of
// the system's own second moment (a point far from the mass, such as the
// centre of a bounding cube stretched by one distant particle, would make
// them large and the differences S - M c c^T of small nodes meaningless).
__global__ void k_quad_ref(const d4* __restrict__ prefix, unsigned N, double* __restrict__ ref) {
	if (threadIdx.x >= 3) return;
	const d4 t = prefix[N - 1];
	const double c = threadIdx.x == 0 ? t.x : (threadIdx.x == 1 ? t.y : t.z);
	ref[threadIdx.x] = t.m > 0.0 ? c / t.m : 0.0;
}

// Traceless quadrupole Qt = 3 Q - tr(Q) I of every internal node about its
// centre of mass, Q = sum m (x - c)(x - c)^T; stored as q0 = (Qxx, Qxy, Qxz,
// Qyy), q1 = Qyz (Qzz = -Qxx - Qyy).  Nodes of at most kQuadDirect particles
// sum directly in FP64, in one pass over offsets from their first particle
// (Q = S' - M c' c'^T with S' and c' about that particle, whose terms are of
// the node's own size); larger nodes use the prefix differences S - M c c^T.
// In FP64 the prefix rounding measured at most 1.2e-6 of M l^2 (l the
// node's longest side) for nodes of 16 particles and 3.5e-7 for 32, on a
// Plummer sphere of 4M particles with outliers at up to 1000 scale radii
// (sequential prefix sums, which round worse than this blocked scan).  Each thread sums its own node, so a
// wavefront takes as long as its largest node: with 256 here the kernel took
// 0.74 ms at 4M particles.  Single-particle leaves have Qt = 0.
constexpr int kQuadDirect = 16;

__global__ __launch_bounds__(kBlock) void k_node_quad(int N, const int2* __restrict__ node_range,
                                                      const float4* __restrict__ body, const d4* __restrict__ prefix,
                                                      const d6* __restrict__ prefix2, const double* __restrict__ ref,
                                                      float4* __restrict__ q0, float* __restrict__ q1) {
	const int n = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (n >= 2 * N - 1) return;
	if (n >= N - 1) {
		q0[n] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		q1[n] = 0.0f;
		return;
	}
	const int2 rg = node_range[n];
	const int f = rg.x, l = rg.x + rg.y - 1;
	const d4 a = f > 0 ? prefix[f - 1] : d4::zero();
	const d6 a2 = f > 0 ? prefix2[f - 1] : d6::zero();
	const d4 b = prefix[l];
	const d6 b2 = prefix2[l];
	const double M = b.m - a.m;
	if (!(M > 0.0)) {
		q0[n] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		q1[n] = 0.0f;
		return;
	}
	double qxx, qxy, qxz, qyy, qyz, qzz;
	if (rg.y <= kQuadDirect) {
		const float4 p0 = body[f];
		double sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
		qxx = qxy = qxz = qyy = qyz = qzz = 0.0;
		for (int k = f; k <= l; ++k) {
			const float4 p = body[k];
			const double m = p.w, x = (double)p.x - (double)p0.x, y = (double)p.y - (double)p0.y,
			             z = (double)p.z - (double)p0.z;
			sm += m;
			sx += m * x;
			sy += m * y;
			sz += m * z;
			qxx += m * x * x;
			qxy += m * x * y;
			qxz += m * x * z;
			qyy += m * y * y;
			qyz += m * y * z;
			qzz += m * z * z;
		}
		const double cx = sx / sm, cy = sy / sm, cz = sz / sm;
		qxx -= sm * cx * cx;
		qxy -= sm * cx * cy;
		qxz -= sm * cx * cz;
		qyy -= sm * cy * cy;
		qyz -= sm * cy * cz;
		qzz -= sm * cz * cz;
	} else {
		// centre of mass relative to ref
		const double cx = (b.x - a.x) / M - ref[0], cy = (b.y - a.y) / M - ref[1], cz = (b.z - a.z) / M - ref[2];
		qxx = (b2.v[0] - a2.v[0]) - M * cx * cx;
		qxy = (b2.v[1] - a2.v[1]) - M * cx * cy;
		qxz = (b2.v[2] - a2.v[2]) - M * cx * cz;
		qyy = (b2.v[3] - a2.v[3]) - M * cy * cy;
		qyz = (b2.v[4] - a2.v[4]) - M * cy * cz;
		qzz = (b2.v[5] - a2.v[5]) - M * cz * cz;
	}
	const double tr = qxx + qyy + qzz;
	q0[n] = make_float4((float)(3.0 * qxx - tr), (float)(3.0 * qxy), (float)(3.0 * qxz), (float)(3.0 * qyy - tr));
	q1[n] = (float)(3.0 * qyz);
	(void)qzz;
}

__device__ __forceinline__ unsigned compact21(unsigned long long v) {
	v &= 0x1249249249249249ull;
	v = (v | (v >> 2)) & 0x10C30C30C30C30C3ull;
	v = (v | (v >> 4)) & 0x100F00F00F00F00Full;
	v = (v | (v >> 8)) & 0x1F0000FF0000FFull;
	v = (v | (v >> 16)) & 0x1F00000000FFFFull;
	v = (v | (v >> 32)) & 0x1FFFFFull;
	return (unsigned)v;
}

// Nodes of at most this many particles take their particles' bounding box as
// their size in the opening test; larger ones their Morton cell.  Each
// thread scans its node's range, so a wavefront takes as long as its largest
// node: 64 kept the build at ~0.15 ms more than Morton cells alone, where 256
// cost ~0.3 ms, for 0.7 % more interactions (a CPU model of the walk).
constexpr int kAabbNode = 64;

// Monopole (node_cm) of every node, and the box of its opening test
// (node_cell: centre, longest side): the bounding box of its particles for
// nodes of at most kAabbNode particles (a point for a leaf), its Morton cell
// above.  For small nodes the particles' box is often much smaller than the
// Morton cell, so at equal accuracy the walk accepts cells sooner (a CPU model
// of the walk at N = 1M: 15 % fewer interactions and 12 % fewer node visits
// for the same 99th-percentile force error).
__global__ __launch_bounds__(kBlock) void k_node_moments(const float4* __restrict__ body, int N,
                                                         const unsigned long long* __restrict__ key,
                                                         const int2* __restrict__ node_range,
                                                         const d4* __restrict__ prefix, const double* __restrict__ box,
                                                         float4* __restrict__ node_cm, float4* __restrict__ node_cell) {
	const int n = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (n >= 2 * N - 1) return;
	if (n >= N - 1) {
		const float4 b = body[n - (N - 1)];
		node_cm[n] = b;
		node_cell[n] = make_float4(b.x, b.y, b.z, 0.0f);
		return;
	}
	const int2 rg = node_range[n];
	const int f = rg.x, l = rg.x + rg.y - 1;
	const d4 hi = prefix[l];
	const d4 lo = f > 0 ? prefix[f - 1] : d4::zero();
	const double M = hi.m - lo.m;
	// cell of the common code prefix: p leading bits of the 63-bit code
	const unsigned long long ka = key[f], kb = key[l];
	const int p = ka == kb ? 63 : __clzll((long long)(ka ^ kb)) - 1;
	const double inv_scale = 1.0 / box[3];
	double centre[3];
	int bmin = 21;
	for (int a = 0; a < 3; ++a) {
		const int bits = max(p - a + 2, 0) / 3;   // code bits fixed on this axis
		bmin = min(bmin, bits);
		const unsigned q = compact21(ka >> (2 - a));
		const unsigned span = 1u << (21 - bits);
		const unsigned q0 = bits == 0 ? 0u : (q & ~(span - 1u));
		centre[a] = box[a] + ((double)q0 + 0.5 * (double)span) * inv_scale;
	}
	const double side = (double)(1u << (21 - bmin)) * inv_scale;
	float4 cm;
	if (M > 0.0) {
		cm = make_float4((float)((hi.x - lo.x) / M), (float)((hi.y - lo.y) / M), (float)((hi.z - lo.z) / M),
		                 (float)M);
	} else {
		cm = make_float4((float)centre[0], (float)centre[1], (float)centre[2], 0.0f);
	}
	node_cm[n] = cm;
	if (rg.y <= kAabbNode) return;   // its particles' box: k_group_boxes
	node_cell[n] = make_float4((float)centre[0], (float)centre[1], (float)centre[2], (float)side);
}

// The opening-test boxes of the internal nodes of at most kAabbNode
// particles: their particles' bounding box (centre, longest side).  These
// are the internal nodes of the walk groups, the maximal subtrees of at most
// 64 particles, and a radix-tree node's index is one end of its range, so a
// group's internal nodes are the indices of its sorted range whose range
// lies inside it.  One wavefront per group: a sparse table of min and max
// over the group's particles (windows of 2^k particles, k = 0 .. 6) in LDS,
// then one lane per internal node combines the two windows that cover its
// range.  Min and max are exact, so each box equals a scan over the node's
// range.
constexpr int kBoxLevels = 7;
constexpr unsigned kBoxWaves = kBlock / 64;
static_assert(kAabbNode == 64, "the walk groups hold at most 64 particles");
__global__ __launch_bounds__(kBlock) void k_group_boxes(const float4* __restrict__ body, int N,
                                                        const unsigned* __restrict__ start,
                                                        const int2* __restrict__ node_range,
                                                        float4* __restrict__ node_cell) {
	// one table at a time (the minima, then the maxima): 7 KiB per wavefront
	__shared__ float s_tab[kBoxWaves][kBoxLevels][3][64];
	const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	const unsigned c = blockIdx.x * kBoxWaves + w;   // chunk of 64 sorted positions
	if (c >= ((unsigned)N + 63u) / 64u) return;
	// the groups that start in the chunk; each ends at the next start flag,
	// at most 64 positions later (start[N] is set)
	const unsigned at = 64u * c + lane;
	const unsigned long long e1 = __ballot(start[at] != 0u);
	const unsigned long long e2 = __ballot(start[at + 64u] != 0u);
	unsigned long long rem = e1 & __ballot(at < (unsigned)N);
	if (!rem) return;
	// the bodies and node ranges of the chunk and of the part of the next one
	// that its last group reaches, loaded once and handed to each group by
	// lane shuffles
	const int over = e2 ? __builtin_ctzll(e2) : 64;
	const int p0 = (int)at, p1 = p0 + 64;
	const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
	const float4 b0 = p0 < N ? body[p0] : zero4;
	const float4 b1 = ((int)lane < over && p1 < N) ? body[p1] : zero4;
	const int2 r0 = p0 < N - 1 ? node_range[p0] : make_int2(0, 0);
	const int2 r1 = ((int)lane < over && p1 < N - 1) ? node_range[p1] : make_int2(0, 0);
	while (rem) {
		const int s0 = __builtin_ctzll(rem);
		rem &= rem - 1ull;
		const unsigned long long above = s0 < 63 ? e1 & (~0ull << (s0 + 1)) : 0ull;
		const int e0 = above ? __builtin_ctzll(above) : (e2 ? 64 + __builtin_ctzll(e2) : 128);
		const int base = 64 * (int)c + s0;
		const int cnt = min(min(e0 - s0, 64), N - base);
		if (cnt < 2) continue;   // a single particle: no internal node
		const bool live = (int)lane < cnt;
		// position base + lane: lane s0 + lane of the chunk's loads, or of the
		// next chunk's.  Every lane takes part in every shuffle (a lane left
		// out of a shuffle does not provide its value to the others), and
		// then picks its half.
		const int src = (s0 + (int)lane) & 63;
		const bool hi_half = s0 + (int)lane >= 64;
		const float x0 = __shfl(b0.x, src), y0 = __shfl(b0.y, src), z0 = __shfl(b0.z, src);
		const float x1 = __shfl(b1.x, src), y1 = __shfl(b1.y, src), z1 = __shfl(b1.z, src);
		const int f0 = __shfl(r0.x, src), n0 = __shfl(r0.y, src);
		const int f1 = __shfl(r1.x, src), n1 = __shfl(r1.y, src);
		const float4 b = hi_half ? make_float4(x1, y1, z1, 0.f) : make_float4(x0, y0, z0, 0.f);
		const int2 rg = hi_half ? make_int2(f1, n1) : make_int2(f0, n0);
		const int idx = base + (int)lane;
		// an internal node of this group: its range lies inside the group
		const bool node = live && idx < N - 1 && rg.x >= base && rg.x + rg.y <= base + cnt;
		const int a0 = rg.x - base, len = max(rg.y, 1);
		const int k = 31 - __clz(len);
		const int a1 = a0 + len - (1 << k);
		float l3[3], h3[3];
#pragma unroll
		for (int pass = 0; pass < 2; ++pass) {
			float v[3] = {b.x, b.y, b.z};
#pragma unroll
			for (int lv = 0; lv < kBoxLevels; ++lv) {
				if (lv > 0) {
					// window [j, j + 2^lv) from [j, j + 2^(lv-1)) and the next one;
					// the windows that run past the group are never read
					const unsigned d = 1u << (lv - 1);
#pragma unroll
					for (int a = 0; a < 3; ++a) {
						const float o = __shfl_down(v[a], d);
						v[a] = pass ? fmaxf(v[a], o) : fminf(v[a], o);
					}
				}
#pragma unroll
				for (int a = 0; a < 3; ++a) s_tab[w][lv][a][lane] = v[a];
			}
			__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
			if (node) {
#pragma unroll
				for (int a = 0; a < 3; ++a) {
					const float x0 = s_tab[w][k][a][a0], x1 = s_tab[w][k][a][a1];
					if (pass) h3[a] = fmaxf(x0, x1);
					else l3[a] = fminf(x0, x1);
				}
			}
			__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		}
		if (node) {
			const float side_box = fmaxf(fmaxf(h3[0] - l3[0], h3[1] - l3[1]), h3[2] - l3[2]);
			node_cell[idx] =
			    make_float4(0.5f * (l3[0] + h3[0]), 0.5f * (l3[1] + h3[1]), 0.5f * (l3[2] + h3[2]), side_box);
		}
	}
}

// Walk records: (r_crit^2, a, b, q) with r_crit = side / theta +
// |com - box centre| (node_cell's box).  Nodes of at most leaf_size particles
// carry (a, b) = (first, count) and the sign bit of the r_crit^2 word set;
// larger ones (left, right).  q is the node's size l = side (float bits),
// which the relative opening test takes to M l^2 and M l^3.
__global__ __launch_bounds__(kBlock) void k_records(int N, const int2* __restrict__ node_lr,
                                                    const int2* __restrict__ node_range,
                                                    const float4* __restrict__ node_cm,
                                                    const float4* __restrict__ node_cell, float inv_theta,
                                                    int leaf_size, int4* __restrict__ rec_b) {
	const int n = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (n >= 2 * N - 1) return;
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
	const float4 cm = node_cm[n], cell = node_cell[n];
	const float ox = cm.x - cell.x, oy = cm.y - cell.y, oz = cm.z - cell.z;
	const float rc = cell.w * inv_theta + sqrtf(ox * ox + oy * oy + oz * oz);
	const bool leafy = count <= leaf_size;
	const int rc2 = __float_as_int(rc * rc) | (leafy ? (int)0x80000000u : 0);
	rec_b[n] = make_int4(rc2, leafy ? first : l, leafy ? count : r, __float_as_int(cell.w));
}

// Walk groups: the maximal subtrees holding at most `cap` particles.  They
// partition the sorted order into runs of at most cap particles that each lie
// inside one Morton cell; start[p] = 1 where a run begins, start[N] = 1 ends
// the last run.  `start` is zero beforehand.
__global__ __launch_bounds__(kBlock) void k_group_flags(int N, const int2* __restrict__ node_lr,
                                                        const int2* __restrict__ node_range, int cap,
                                                        unsigned* __restrict__ start) {
	const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
	if (i == 0) {
		start[N] = 1u;
		if (N <= cap) start[0] = 1u;   // the root (a single leaf when N == 1) is the only group
	}
	if (i >= N - 1) return;
	if (node_range[i].y <= cap) return;
	const int2 lr = node_lr[i];
	const int child[2] = {lr.x, lr.y};
	for (int k = 0; k < 2; ++k) {
		const int c = child[k];
		const int2 rg = c >= N - 1 ? make_int2(c - (N - 1), 1) : node_range[c];
		if (rg.y <= cap) start[rg.x] = 1u;
	}
}

}  // namespace

extern "C" {

unsigned sfd_lbvh_group_flag_len(unsigned N) { return (N + 63u) / 64u * 64u + 64u; }

int sfd_lbvh_groups(unsigned N, const void* node_lr, const void* node_range, int cap, unsigned* start,
                    hipStream_t s) {
	if (N == 0 || cap < 1 || cap > 64) return -1;
	const unsigned len = sfd_lbvh_group_flag_len(N);
	hipLaunchKernelGGL(k_fill_u32, dim3((len + kBlock - 1) / kBlock), dim3(kBlock), 0, s, start, len, 0u);
	const unsigned threads = N > 1 ? N - 1 : 1;
	hipLaunchKernelGGL(k_group_flags, dim3((threads + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (int)N,
	                   static_cast<const int2*>(node_lr), static_cast<const int2*>(node_range), cap, start);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_groups", "launch");
}

int sfd_lbvh_group_boxes(unsigned N, const void* body, const unsigned* start, const void* node_range,
                         void* node_cell, hipStream_t s) {
	if (N < 2) return 0;
	const unsigned chunks = (N + 63u) / 64u;
	hipLaunchKernelGGL(k_group_boxes, dim3((chunks + kBoxWaves - 1) / kBoxWaves), dim3(kBlock), 0, s,
	                   static_cast<const float4*>(body), (int)N, start, static_cast<const int2*>(node_range),
	                   static_cast<float4*>(node_cell));
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_group_boxes", "launch");
}

unsigned sfd_lbvh_bbox_blocks(void) { return 256u; }

int sfd_lbvh_keys(const float* pos, const float* mas, unsigned N, unsigned n_sort, double* bbox_part, double* box,
                  unsigned long long* key, unsigned* val, void* packed, hipStream_t s) {
	const unsigned nb = sfd_lbvh_bbox_blocks();
	if (N == 0 || n_sort % kTile != 0) return -1;
	hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kBlock), 0, s, pos, N, bbox_part);
	hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, s, bbox_part, nb, box);
	hipLaunchKernelGGL(k_morton, dim3((n_sort + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, mas, N, n_sort, box,
	                   key, val, static_cast<float4*>(packed));
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_keys", "launch");
}

unsigned sfd_radix_tile(void) { return kTile; }

unsigned sfd_radix_hist_entries(unsigned n_sort) {
	const unsigned nh = 256u * (n_sort / kTile);
	return nh + (nh + kScanU32Chunk - 1u) / kScanU32Chunk;
}

unsigned sfd_lbvh_scan_tiles(unsigned N) { return (N + kScanTile - 1) / kScanTile; }

// Histograms of up to this many entries (128 tiles, 524288 keys) are scanned
// by one block in one launch instead of three.
constexpr unsigned kRadixScanOneBlock = 32768;

// Sorts n_sort (a multiple of the tile) pairs on the low 8 * passes key bits;
// the result ends in (key0, val0) after an even number of passes and in
// (key1, val1) after an odd one.
int sfd_radix_sort_passes(unsigned long long* key0, unsigned* val0, unsigned long long* key1, unsigned* val1,
                          unsigned n_sort, unsigned* hist, unsigned passes, hipStream_t s) {
	if (n_sort % kTile != 0 || passes > 8u) return -1;
	const unsigned ntiles = n_sort / kTile;
	const unsigned nh = 256u * ntiles;
	const unsigned nchunks = (nh + kScanU32Chunk - 1u) / kScanU32Chunk;
	unsigned* chunk_sums = hist + nh;   // sfd_radix_hist_entries() leaves room for them
	if (ntiles == 0) return 0;
	for (unsigned pass = 0; pass < passes; ++pass) {
		const unsigned shift = 8u * pass;
		unsigned long long* kin = (pass & 1) ? key1 : key0;
		unsigned* vin = (pass & 1) ? val1 : val0;
		unsigned long long* kout = (pass & 1) ? key0 : key1;
		unsigned* vout = (pass & 1) ? val0 : val1;
		hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, s, kin, shift, ntiles, hist);
		if (nh <= kRadixScanOneBlock) {
			hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, hist, nh);
		} else {
			hipLaunchKernelGGL(k_scan_u32_sums, dim3(nchunks), dim3(kBlock), 0, s, hist, nh, chunk_sums);
			hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, chunk_sums, nchunks);
			hipLaunchKernelGGL(k_scan_u32_down, dim3(nchunks), dim3(kBlock), 0, s, hist, nh, chunk_sums);
		}
		hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kBlock), 0, s, kin, vin, shift, ntiles, hist, n_sort,
		                   kout, vout);
	}
	return sf_hip_check(hipGetLastError(), "sfd_radix_sort", "launch");
}

int sfd_radix_sort(unsigned long long* key0, unsigned* val0, unsigned long long* key1, unsigned* val1,
                   unsigned n_sort, unsigned* hist, hipStream_t s) {
	return sfd_radix_sort_passes(key0, val0, key1, val1, n_sort, hist, 8u, s);
}

unsigned sfd_scan_u32_sums_len(unsigned n) { return (n + kScanU32Chunk - 1u) / kScanU32Chunk; }

// In-place exclusive scan of n u32 entries; sums holds
// sfd_scan_u32_sums_len(n) entries of scratch.
int sfd_scan_u32_exclusive(unsigned* data, unsigned n, unsigned* sums, hipStream_t s) {
	if (n == 0) return 0;
	const unsigned nchunks = sfd_scan_u32_sums_len(n);
	hipLaunchKernelGGL(k_scan_u32_sums, dim3(nchunks), dim3(kBlock), 0, s, data, n, sums);
	hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, sums, nchunks);
	hipLaunchKernelGGL(k_scan_u32_down, dim3(nchunks), dim3(kBlock), 0, s, data, n, sums);
	return sf_hip_check(hipGetLastError(), "sfd_scan_u32_exclusive", "launch");
}

// prefix: N records of 4 doubles; tile_sum: sfd_lbvh_scan_tiles(N) records
int sfd_lbvh_topology(const unsigned long long* key, const unsigned* order, unsigned N, const double* box, void* body,
                      void* node_lr, void* node_range, void* prefix, void* tile_sum, void* node_cm, void* node_cell,
                      hipStream_t s) {
	const unsigned g = (N + kBlock - 1) / kBlock;
	float4* b4 = static_cast<float4*>(body);
	int2* rg = static_cast<int2*>(node_range);
	d4* pre = static_cast<d4*>(prefix);
	d4* ts = static_cast<d4*>(tile_sum);
	const unsigned ntiles = sfd_lbvh_scan_tiles(N);
	const int tok = sf_prof_begin("lbvh_topology", s);
	// the packed records of sfd_lbvh_keys, in the prefix buffer until the scan overwrites it
	hipLaunchKernelGGL(k_gather_sorted, dim3(g), dim3(kBlock), 0, s, reinterpret_cast<const float4*>(prefix), order, N,
	                   b4);
	if (N > 1)
		hipLaunchKernelGGL(k_karras, dim3((N - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, key, (int)N,
		                   static_cast<int2*>(node_lr), rg);
	hipLaunchKernelGGL(k_scan_tiles<d4>, dim3(ntiles), dim3(kBlock), 0, s, b4, N, (const double*)nullptr, ts);
	hipLaunchKernelGGL(k_scan_tile_sums<d4>, dim3(1), dim3(kBlock), 0, s, ts, ntiles);
	hipLaunchKernelGGL(k_scan_write<d4>, dim3(ntiles), dim3(kBlock), 0, s, b4, N, (const double*)nullptr, ts, pre);
	const unsigned nn = 2u * N - 1u;
	hipLaunchKernelGGL(k_node_moments, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b4, (int)N, key, rg, pre,
	                   box, static_cast<float4*>(node_cm), static_cast<float4*>(node_cell));
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_topology", "launch");
}

// Traceless quadrupoles of every node (after sfd_lbvh_topology): prefix2 is
// N records of 6 doubles, tile_sum2 sfd_lbvh_scan_tiles(N) of them, ref 3
// doubles; node_q0 / node_q1 are 2N-1 float4 / float.
int sfd_lbvh_quadrupoles(unsigned N, const void* body, const void* node_range, const void* prefix, const double* box,
                         void* prefix2, void* tile_sum2, double* ref, void* node_q0, float* node_q1, hipStream_t s) {
	(void)box;
	if (N == 0) return 0;
	const float4* b4 = static_cast<const float4*>(body);
	d6* pre2 = static_cast<d6*>(prefix2);
	d6* ts2 = static_cast<d6*>(tile_sum2);
	const unsigned ntiles = sfd_lbvh_scan_tiles(N);
	hipLaunchKernelGGL(k_quad_ref, dim3(1), dim3(64), 0, s, static_cast<const d4*>(prefix), N, ref);
	hipLaunchKernelGGL(k_scan_tiles<d6>, dim3(ntiles), dim3(kBlock), 0, s, b4, N, (const double*)ref, ts2);
	hipLaunchKernelGGL(k_scan_tile_sums<d6>, dim3(1), dim3(kBlock), 0, s, ts2, ntiles);
	hipLaunchKernelGGL(k_scan_write<d6>, dim3(ntiles), dim3(kBlock), 0, s, b4, N, (const double*)ref, ts2, pre2);
	const unsigned nn = 2u * N - 1u;
	hipLaunchKernelGGL(k_node_quad, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (int)N,
	                   static_cast<const int2*>(node_range), b4, static_cast<const d4*>(prefix), pre2,
	                   (const double*)ref, static_cast<float4*>(node_q0), node_q1);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_quadrupoles", "launch");
}

int sfd_lbvh_records(unsigned N, const void* node_lr, const void* node_range, const void* node_cm,
                     const void* node_cell, float theta, int leaf_size, void* rec_b, hipStream_t s) {
	const unsigned nn = 2u * N - 1u;
	const float inv_theta = theta > 0.0f ? 1.0f / theta : 3.0e38f;
	hipLaunchKernelGGL(k_records, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (int)N,
	                   static_cast<const int2*>(node_lr), static_cast<const int2*>(node_range),
	                   static_cast<const float4*>(node_cm), static_cast<const float4*>(node_cell), inv_theta,
	                   leaf_size, static_cast<int4*>(rec_b));
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_records", "launch");
}

}  // extern "C"
