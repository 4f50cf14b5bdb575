// lbvh_walk.hip -- the SF_TREE_LBVH walk: one wavefront per walk group, a
// maximal subtree of at most 64 particles (a run of consecutive sorted
// particles inside one Morton cell).  Groups are not fixed runs of 64 sorted
// particles: a run that crosses a high-level Morton boundary (the octant
// planes through the centre of a cluster) has a bounding box the size of the
// whole core and would open nearly every node.
//
// Each step pops up to 64 pending nodes, one per lane, and classifies them
// against the group's bounding box: a monopole "cell" when the box stays
// outside the node's critical radius r = l/theta + |com - centre| (l = longest
// side of the node's particle box), a "leaf" summed particle by particle when
// the node holds at most leaf_size particles, otherwise opened (its two
// children are pushed).  Cells and leaf particles are queued in LDS and
// evaluated 64 at a time by every lane with the PP kernel's 13-VALU + rsqrt
// interaction; leaf particles are gathered 64 at a time with one vector load,
// so the tree logic and memory latency are amortised over 64 interactions.
//
// Options: quadrupole cells (the 34-VALU interact_quad), and the relative
// opening test, which also requires G M l^p <= alpha |a|min d^(p+2) of a
// cell (|a|min the group's smallest previous acceleration); a quadrupole
// walk then sums the cells that meet the monopole form of that bound (scaled
// by kMonoBeta) as monopoles, through the leaf particles' queue.
//
// The grid is persistent and sized to the wavefronts resident at once; each
// wavefront takes runs of 64 sorted positions from a shared counter, walks
// the groups that start in each, and owns a node stack: its top in LDS, the
// rest in global memory (deep enough that the 64-wide pops rarely have to
// narrow).  The self pair is included in leaf sums, as in the reference's PP
// sum.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"
#include "tree_dev.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kList = 128;          // queued cells / leaves / particle ids per wavefront
constexpr int kStack = 8192;        // node ids per wavefront (global memory)
constexpr int kHead = 160;          // room for a depth-first descent (tree depth <= 96)
// The top of each wavefront's node stack lives in LDS: a batch's children
// are pushed there and the next batch pops them without a round trip to L2.
// Children that do not fit go to the global stack, which is popped when the
// LDS top holds fewer than a batch.  The quadrupole walk's LDS (its cells
// carry 24 more bytes) leaves room for a smaller top at 5 wavefronts per SIMD.
constexpr int kTopMono = 256;
constexpr int kTopQuad = 128;
// A quadrupole walk under the relative opening test takes a cell as a
// monopole when G M l^2 <= kMonoBeta alpha |a| d^4.  The factor calibrates
// the two size terms against each other.  Swept at 4M particles (quadrupole
// walk, theta 1.15, alpha 2e-3 to 4e-3): up to 0.2 the 99th-percentile force
// error rises by under 3 % while the walk gets 18 % faster; 0.5 raises it by
// half, 1 more than quadruples it.
constexpr float kMonoBeta = 0.2f;

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__device__ __forceinline__ float wave_min(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
	return v;
}
__device__ __forceinline__ float wave_max(float v) {
	for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
	return v;
}

struct Acc {
	float tx, ty, tz, tp;   // partial sums of the current batch of 64
	float sx, sy, sz, sp;   // running sums
	float cx, cy, cz, cp;   // Kahan compensations
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
	const float y = x - c;
	const float t = s + y;
	c = (t - s) - y;
	s = t;
}

__device__ __forceinline__ void fold(Acc& a) {
	kahan_add(a.sx, a.cx, a.tx);
	kahan_add(a.sy, a.cy, a.ty);
	kahan_add(a.sz, a.cz, a.tz);
	kahan_add(a.sp, a.cp, a.tp);
	a.tx = a.ty = a.tz = a.tp = 0.0f;
}

// UNIFORM: every q has the same mass, left out of the sums (the caller
// scales the partial by it before the fold)
template <bool UNIFORM = false>
__device__ __forceinline__ void interact(Acc& a, float xi, float yi, float zi, float4 bj, float eps2) {
	const float dx = bj.x - xi, dy = bj.y - yi, dz = bj.z - zi;
	float r2 = __builtin_fmaf(dx, dx, eps2);
	r2 = __builtin_fmaf(dy, dy, r2);
	r2 = __builtin_fmaf(dz, dz, r2);
	const float inv = __builtin_amdgcn_rsqf(r2);
	const float mi = UNIFORM ? inv : bj.w * inv;
	const float w = mi * (inv * inv);
	a.tx = __builtin_fmaf(w, dx, a.tx);
	a.ty = __builtin_fmaf(w, dy, a.ty);
	a.tz = __builtin_fmaf(w, dz, a.tz);
	a.tp += mi;
}

// Interactions with q[0 .. n), n <= 64, read from LDS by every lane.
template <bool UNIFORM = false>
__device__ __forceinline__ void eval_list(Acc& a, float xi, float yi, float zi, const float4* q, int n, float eps2,
                                          float m = 1.0f) {
	if (n == 64) {
#pragma unroll 8
		for (int k = 0; k < 64; ++k) interact<UNIFORM>(a, xi, yi, zi, q[k], eps2);
	} else {
		for (int k = 0; k < n; ++k) interact<UNIFORM>(a, xi, yi, zi, q[k], eps2);
	}
	if (UNIFORM) {
		a.tx *= m;
		a.ty *= m;
		a.tz *= m;
		a.tp *= m;
	}
	fold(a);
}

// Monopole plus traceless quadrupole of a cell with e = com - x:
//   acc += M e / r^3 - Qt e / r^5 + (5/2) (e^T Qt e) e / r^7
//   pot += M / r + (1/2) (e^T Qt e) / r^5
// (the potential -G [M/r + (1/2) d^T Qt d / r^5] with d = -e and its
// gradient), Plummer-softened through r^2 = |e|^2 + eps^2.  Qt is
// (q0 = Qxx, Qxy, Qxz, Qyy; q1 = Qyz, Qzz), Qzz = -Qxx - Qyy stored with the
// queued cell.  34 VALU operations plus the rsqrt.
__device__ __forceinline__ void interact_quad(Acc& a, float xi, float yi, float zi, float4 c, float4 q0, float2 q1,
                                              float eps2) {
	const float ex = c.x - xi, ey = c.y - yi, ez = c.z - zi;
	float r2 = __builtin_fmaf(ex, ex, eps2);
	r2 = __builtin_fmaf(ey, ey, r2);
	r2 = __builtin_fmaf(ez, ez, r2);
	const float inv = __builtin_amdgcn_rsqf(r2);
	const float inv2 = inv * inv;
	const float mi = c.w * inv;
	const float m3 = mi * inv2;
	const float inv5 = (inv * inv2) * inv2;
	const float qx = __builtin_fmaf(q0.x, ex, __builtin_fmaf(q0.y, ey, q0.z * ez));
	const float qy = __builtin_fmaf(q0.y, ex, __builtin_fmaf(q0.w, ey, q1.x * ez));
	const float qz = __builtin_fmaf(q0.z, ex, __builtin_fmaf(q1.x, ey, q1.y * ez));
	const float eqe = __builtin_fmaf(ex, qx, __builtin_fmaf(ey, qy, ez * qz));
	const float t = eqe * inv5;
	const float w = __builtin_fmaf(2.5f, t * inv2, m3);
	a.tx = __builtin_fmaf(w, ex, __builtin_fmaf(-inv5, qx, a.tx));
	a.ty = __builtin_fmaf(w, ey, __builtin_fmaf(-inv5, qy, a.ty));
	a.tz = __builtin_fmaf(w, ez, __builtin_fmaf(-inv5, qz, a.tz));
	a.tp = __builtin_fmaf(0.5f, t, a.tp) + mi;
}

__device__ __forceinline__ void eval_list_quad(Acc& a, float xi, float yi, float zi, const float4* c, const float4* q0,
                                               const float2* q1, int n, float eps2) {
	if (n == 64) {
#pragma unroll 4
		for (int k = 0; k < 64; ++k) interact_quad(a, xi, yi, zi, c[k], q0[k], q1[k], eps2);
	} else {
		for (int k = 0; k < n; ++k) interact_quad(a, xi, yi, zi, c[k], q0[k], q1[k], eps2);
	}
	fold(a);
}

// 5 wavefronts per SIMD (96 VGPRs): the monopole walk fits without spills
// and runs 7% faster at 4M than at 4 per SIMD; the quadrupole walk spills 11
// VGPRs, and at 4 per SIMD without spills it was 3.4% slower.
// QUAD: cells carry traceless quadrupoles (node_q0, node_q1).
template <bool QUAD>
__global__ __launch_bounds__(kBlock, 5) void k_walk(const float4* __restrict__ body, const unsigned* __restrict__ order,
                                                 int N, const unsigned* __restrict__ group_start,
                                                 const float4* __restrict__ rec_a,
                                                 const float4* __restrict__ node_q0, const float* __restrict__ node_q1,
                                                 const int4* __restrict__ rec_b, float alpha_g, float eps2, double G, int acc_phys,
                                                 float* __restrict__ acc, float* __restrict__ pot,
                                                 unsigned long long* __restrict__ counters, int* __restrict__ gstack,
                                                 unsigned n_chunks) {
	__shared__ float4 s_cell[kWaves][kList];
	__shared__ int2 s_leaf[kWaves][kList];
	__shared__ int s_pid[kWaves][kList];
	__shared__ float4 s_part[kWaves][64];
	__shared__ float4 s_q0[kWaves][QUAD ? kList : 1];
	__shared__ float2 s_q1[kWaves][QUAD ? kList : 1];
	constexpr int kTop = QUAD ? kTopQuad : kTopMono;
	__shared__ int s_top[kWaves][kTop];
	const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	float4* cq0 = s_q0[w];
	float2* cq1 = s_q1[w];
	int* stack = gstack + (size_t)(blockIdx.x * kWaves + w) * kStack;
	int* top = s_top[w];
	float4* cells = s_cell[w];
	int2* leaves = s_leaf[w];
	int* pid = s_pid[w];
	float4* part = s_part[w];
	const unsigned long long lt = (1ull << lane) - 1ull;
	const int n_nodes = 2 * N - 1;
	const float m0 = body[0].w;   // the mass of the equal-mass leaf batches
	const unsigned budget = 2u * (unsigned)n_nodes + 64u;
	unsigned long long n_cell_tot = 0, n_part_tot = 0, n_batch = 0;
	unsigned bad = 0;

	// chunks of 64 sorted positions are taken from a shared counter, one
	// atomic per chunk: walk groups differ in cost (more so under the
	// relative opening test), and a static split leaves waves idle at the end
	for (;;) {
		unsigned long long cw = 0;
		if (lane == 0) cw = atomicAdd(&counters[5], 1ull);
		const unsigned c = (unsigned)__shfl(cw, 0);
		if (c >= n_chunks) break;
		// the groups that start in sorted positions [64c, 64c + 64); each ends at
		// the next start flag, at most 64 positions later (start[N] is set)
		const unsigned at = 64u * c + lane;
		const unsigned long long e1 = __ballot(group_start[at] != 0u);
		const unsigned long long e2 = __ballot(group_start[at + 64u] != 0u);
		unsigned long long rem = e1 & __ballot(at < (unsigned)N);
		while (rem) {
			const int s0 = __builtin_ctzll(rem);
			rem &= rem - 1ull;
			const unsigned long long above = s0 < 63 ? e1 & (~0ull << (s0 + 1)) : 0ull;
			const int e0 = above ? __builtin_ctzll(above) : (e2 ? 64 + __builtin_ctzll(e2) : 128);
			const int base = 64 * (int)c + s0;
			int cnt = e0 - s0;
			if (cnt > 64 || cnt > N - base) {   // only a malformed start array gets here
				bad = 1u;
				cnt = min(64, N - base);
			}
			const int p = base + (int)lane;
			const bool live = (int)lane < cnt;
			const float4 me = body[live ? p : base];
			const float xi = me.x, yi = me.y, zi = me.z;
			const float gx0 = wave_min(xi), gy0 = wave_min(yi), gz0 = wave_min(zi);
			const float gx1 = wave_max(xi), gy1 = wave_max(yi), gz1 = wave_max(zi);
			// relative opening test: am = alpha / G times the smallest previous
			// |a| of the group (0: theta test alone; also when no particle has a
			// positive finite previous acceleration)
			float am = 0.0f;
			if (alpha_g > 0.0f) {
				float amag = 3.0e38f;
				if (live) {
					const unsigned i0 = order[p];
					const float ax = acc[3u * i0], ay = acc[3u * i0 + 1u], az = acc[3u * i0 + 2u];
					const float a = sqrtf(ax * ax + ay * ay + az * az);
					amag = acc_phys ? a : a / me.w;   // the reference's acc is m a
				}
				amag = wave_min(amag);   // fminf drops NaN
				am = (amag > 0.0f && amag < 3.0e38f) ? alpha_g * amag : 0.0f;
			}

			Acc a;
			a.tx = a.ty = a.tz = a.tp = a.sx = a.sy = a.sz = a.sp = a.cx = a.cy = a.cz = a.cp = 0.0f;
			unsigned long long n_cell = 0, n_part = 0, n_mono = 0;
			unsigned popped = 0;
			// pending nodes: nt in the LDS top, ng in the global stack
			int nt = 1, ng = 0, nc = 0, nl = 0, npid = 0;
			if (lane == 0) top[0] = 0;   // root: internal node 0, or the only leaf when N == 1
			wave_fence();

			while (nt + ng > 0 || nc > 0 || nl > 0 || npid > 0) {
				if (nt + ng > 0) {
					// pops come from the LDS top first; the global stack grows by at
					// most 2 k per batch, and by one per batch in the narrow regime
					const int k = (ng + 128 <= kStack - kHead) ? min(64, nt + ng) : 1;
					const int kt = min(k, nt);
					int node = -1;
					if ((int)lane < kt) node = top[nt - 1 - (int)lane];
					else if ((int)lane < k) node = stack[ng - 1 - ((int)lane - kt)];
					nt -= kt;
					ng -= k - kt;
					popped += (unsigned)k;
					++n_batch;
					wave_fence();
					if (node >= n_nodes) {
						node = -1;   // only a malformed tree gets here
						bad = 1u;
					}
					int cls = 0;   // 0 none, 1 cell, 2 leaf, 3 open, 4 monopole of a quadrupole walk
					float4 ra = make_float4(0.f, 0.f, 0.f, 0.f);
					int4 rb = make_int4(0, 0, 0, 0);
					if (node >= 0) {
						ra = rec_a[node];
						rb = rec_b[node];
						const float dx = fmaxf(fmaxf(gx0 - ra.x, ra.x - gx1), 0.0f);
						const float dy = fmaxf(fmaxf(gy0 - ra.y, ra.y - gy1), 0.0f);
						const float dz = fmaxf(fmaxf(gz0 - ra.z, ra.z - gz1), 0.0f);
						const float d2 = dx * dx + dy * dy + dz * dz;
						// the sign bit of the r_crit^2 word marks a (first, count) record
						const bool leafy = rb.x < 0;
						// relative test: G M l^p <= alpha |a| d^(p+2), p = 2 for a
						// monopole and 3 for a quadrupole (l = the record's size).  A
						// quadrupole walk takes a cell that meets the p = 2 bound
						// (scaled by kMonoBeta) as a monopole, for 13 VALU instead of 34.
						bool rel = true, mono = false;
						if (am > 0.0f) {
							const float l = __int_as_float(rb.w), d4 = d2 * d2;
							const float ml2 = ra.w * l * l;
							mono = ml2 <= (QUAD ? kMonoBeta : 1.0f) * am * d4;
							rel = QUAD ? ml2 * l <= am * d4 * sqrtf(d2) : mono;
						}
						// a cell holds at least two particles (every internal node does)
						if (d2 > __int_as_float(rb.x & 0x7fffffff) && (!leafy || rb.z > 1) && rel)
							cls = (QUAD && mono) ? 4 : 1;
						else if (leafy) cls = 2;
						else cls = 3;
					}
					const unsigned long long m_cell = __ballot(cls == 1);
					const unsigned long long m_leaf = __ballot(cls == 2);
					const unsigned long long m_open = __ballot(cls == 3);
					// monopoles of a quadrupole walk join the particle queue as
					// negative ids (npid < 64 here, so 64 more fit in pid[])
					const unsigned long long m_mono = QUAD ? __ballot(cls == 4) : 0ull;
					if (QUAD && cls == 4) pid[npid + __popcll(m_mono & lt)] = ~node;
					if (cls == 1) {
						const int at = nc + __popcll(m_cell & lt);
						cells[at] = ra;
						if (QUAD) {
							const float4 q0 = node_q0[node];
							cq0[at] = q0;
							cq1[at] = make_float2(node_q1[node], -q0.x - q0.w);
						}
					}
					if (cls == 2) leaves[nl + __popcll(m_leaf & lt)] = make_int2(rb.y, rb.z);
					if (cls == 3) {
						// children fill the LDS top; what does not fit goes to the
						// global stack
						const int at = nt + 2 * __popcll(m_open & lt);
#pragma unroll
						for (int h = 0; h < 2; ++h) {
							const int child = h ? rb.z : rb.y;
							if (at + h < kTop) {
								top[at + h] = child;
							} else {
								const int gi = ng + at + h - kTop;
								if (gi < kStack) stack[gi] = child;   // always: see kHead
								else bad = 1u;
							}
						}
					}
					nc += __popcll(m_cell);
					nl += __popcll(m_leaf);
					npid += __popcll(m_mono);
					n_mono += (unsigned long long)__popcll(m_mono);
					const int pushed = nt + 2 * __popcll(m_open);
					ng = min(ng + max(pushed - kTop, 0), kStack);
					nt = min(pushed, kTop);
					wave_fence();
					if (popped > budget) {   // a tree visits each node at most once per group
						bad = 1u;
						nt = ng = 0;
					}
				}
				const bool drain = nt + ng == 0;
				// monopoles
				if (nc >= 64 || (drain && nc > 0)) {
					const int take = min(nc, 64);
					if (QUAD) eval_list_quad(a, xi, yi, zi, cells, cq0, cq1, take, eps2);
					else eval_list(a, xi, yi, zi, cells, take, eps2);
					n_cell += (unsigned long long)take;
					wave_fence();
					if ((int)lane < nc - take) {
						cells[lane] = cells[take + lane];
						if (QUAD) {
							cq0[lane] = cq0[take + lane];
							cq1[lane] = cq1[take + lane];
						}
					}
					nc -= take;
					wave_fence();
				}
				// leaves: expand into particle ids, then gather and sum 64 at a time
				if (nl >= 64 || npid >= 64 || (drain && (nl > 0 || npid > 0))) {
					int q = 0;
					while (q < nl || npid >= 64 || (drain && npid > 0)) {
						if (q < nl && npid < 64) {
							// up to 64 leaves at once, one per lane: an inclusive scan of
							// their counts places each leaf's particle ids after the
							// previous ones; the leaves that fit in pid[] (a prefix, the
							// scan being nondecreasing, and at least the first: counts
							// are at most 64 and npid < 64) are taken
							const int li = q + (int)lane;
							const int2 lf = li < nl ? leaves[li] : make_int2(0, 0);
							int incl = lf.y;
#pragma unroll
							for (int o = 1; o < 64; o <<= 1) {
								const int v = __shfl_up(incl, (unsigned)o);
								if ((int)lane >= o) incl += v;
							}
							const bool fits = li < nl && npid + incl <= kList;
							const int k = __popcll(__ballot(fits));
							const int total = __shfl(incl, k - 1);
							if (fits) {
								const int at = npid + incl - lf.y;
								for (int j = 0; j < lf.y; ++j) pid[at + j] = lf.x + j;
							}
							npid += total;
							q += k;
						}
						wave_fence();
						if (npid >= 64 || (drain && q == nl && npid > 0)) {
							const int take = min(npid, 64);
							float pm = m0;
							if ((int)lane < take) {
								const int id = pid[lane];
								const float4 pb = (!QUAD || id >= 0) ? body[id] : rec_a[~id];
								part[lane] = pb;
								pm = pb.w;
							}
							wave_fence();
							// particles of one mass (the common case) leave the
							// per-pair mass product out
							if (__ballot(__float_as_uint(pm) != __float_as_uint(m0)) == 0ull)
								eval_list<true>(a, xi, yi, zi, part, take, eps2, m0);
							else
								eval_list(a, xi, yi, zi, part, take, eps2);
							n_part += (unsigned long long)take;
							wave_fence();
							if ((int)lane < npid - take) pid[lane] = pid[take + (int)lane];
							npid -= take;
							wave_fence();
						} else {
							break;
						}
					}
					wave_fence();
					if ((int)lane < nl - q) leaves[lane] = leaves[q + (int)lane];
					nl -= q;
					wave_fence();
				}
			}
			if (live) {
				const unsigned i = order[p];
				const double sx = (double)a.sx - (double)a.cx, sy = (double)a.sy - (double)a.cy,
				             sz = (double)a.sz - (double)a.cz, sp2 = (double)a.sp - (double)a.cp;
				const double f = acc_phys ? G : G * (double)me.w;
				acc[3u * i + 0u] = (float)(f * sx);
				acc[3u * i + 1u] = (float)(f * sy);
				acc[3u * i + 2u] = (float)(f * sz);
				pot[i] = (float)(-G * sp2);
			}
			n_cell_tot += (n_cell + n_mono) * (unsigned long long)cnt;
			n_part_tot += (n_part - n_mono) * (unsigned long long)cnt;
		}
	}
	const bool any_bad = __ballot(bad != 0u) != 0ull;
	if (counters && lane == 0) {
		atomicAdd(&counters[0], n_cell_tot);
		atomicAdd(&counters[1], n_part_tot);
		if (any_bad) {
			atomicAdd(&counters[2], 1ull);
			atomicAdd(&counters[4], 1ull);   // sticky: cleared when the host checks it
		}
		atomicAdd(&counters[3], n_batch);
	}
}

__global__ void k_walk_reset(unsigned long long* __restrict__ counters) {
	if (threadIdx.x < 6 && threadIdx.x != 4) counters[threadIdx.x] = 0ull;
}

// Blocks that are resident at once, the size of the persistent grid: every
// wavefront runs from the start and takes chunks until none are left
// (cached after the first call).
template <bool QUAD>
unsigned resident_blocks_of(void) {
	static unsigned cached = 0;
	if (cached) return cached;
	int per_cu = 0, cus = 0, dev = 0;
	if (hipGetDevice(&dev) != hipSuccess ||
	    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_walk<QUAD>, kBlock, 0) != hipSuccess ||
	    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1 ||
	    cus < 1) {
		(void)hipGetLastError();
		per_cu = 4;
		cus = 256;
	}
	cached = (unsigned)per_cu * (unsigned)cus;
	return cached;
}

// the node stacks are sized for the larger of the two variants' grids
unsigned resident_blocks(void) { return max(resident_blocks_of<false>(), resident_blocks_of<true>()); }

}  // namespace

extern "C" {

size_t sfd_lbvh_walk_stack_bytes(void) { return (size_t)resident_blocks() * kWaves * kStack * sizeof(int); }

int sfd_lbvh_walk(const void* body, const unsigned* order, unsigned N, const unsigned* group_start,
                  const void* node_cm, const void* node_q0, const float* node_q1, const void* rec_b,
                  float alpha_g, float eps2, double G, int acc_phys, float* acc, float* pot,
                  unsigned long long* counters, int* gstack, hipStream_t s) {
	if (N == 0) return 0;
	if (!counters || !gstack || !group_start) return -1;
	const bool quad = node_q0 && node_q1;
	const unsigned chunks = (N + 63) / 64;   // group_start holds 64 * (chunks + 1) flags
	unsigned blocks = (chunks + kWaves - 1) / kWaves;
	const unsigned resident = quad ? resident_blocks_of<true>() : resident_blocks_of<false>();
	if (blocks > resident) blocks = resident;   // gstack holds resident_blocks() >= resident slots
	hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, s, counters);
	const int tok = sf_prof_begin("lbvh_walk", s);
	if (quad)
		hipLaunchKernelGGL(k_walk<true>, dim3(blocks), dim3(kBlock), 0, s, static_cast<const float4*>(body), order,
		                   (int)N, group_start, static_cast<const float4*>(node_cm),
		                   static_cast<const float4*>(node_q0), node_q1, static_cast<const int4*>(rec_b),
		                   alpha_g, eps2, G, acc_phys, acc, pot, counters, gstack, chunks);
	else
		hipLaunchKernelGGL(k_walk<false>, dim3(blocks), dim3(kBlock), 0, s, static_cast<const float4*>(body), order,
		                   (int)N, group_start, static_cast<const float4*>(node_cm), (const float4*)nullptr,
		                   (const float*)nullptr, static_cast<const int4*>(rec_b), alpha_g, eps2, G, acc_phys,
		                   acc, pot, counters, gstack, chunks);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_lbvh_walk", "launch");
}

}  // extern "C"
