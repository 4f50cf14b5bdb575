// tree_ref_par.hip -- the reference's octree (src/tree.c:120-296) built level
// by level in parallel, node for node and bit for bit equal to the serial
// insertion of tree_ref.hip.
//
// Why the serial insertion can be replayed one depth at a time:
//  * particles reach a node in index order (each insertion ends before the
//    next begins, and a split hands its eight bodies down in index order),
//    so a node's body count, leaf list and FP32 moment sums depend only on
//    its member set taken in index order;
//  * a node splits when its ninth member arrives.  Its first eight members
//    are redistributed with '<' (src/tree.c:244-247), the ninth and every
//    later member descends with '<=' (src/tree.c:282-285): the rule follows
//    from the member's rank at the node;
//  * node X splits while its rank-8 member i9 is inserted, at X's depth, so
//    the serial order of the splits is the order of (i9, depth), and the
//    k-th split (from 0) creates nodes 1 + 8k ... 1 + 8k + 7.
// Each level holds the particles still descending as (slot, index) pairs
// sorted with the stable radix sort; a slot's members are counted, their
// mass-moment products stored in that order, and the members of splitting
// slots routed by rank to their child slot.  Nodes are numbered level-major
// while building.  Every node's moments are summed in index order,
// sequentially as the reference adds them, on a second stream as soon as the
// node's level is finished (the root's chain, the longest, beside the
// deeper levels); the split events, sorted by (i9, depth), map the
// level-major numbering onto the reference's.
//
// The level sizes live on the device (st[]), so the host queues several
// levels at a time and synchronises once per batch.  A build that would run
// out of nodes or past depth_max is reported to the caller, which runs the
// serial build: that one stops where the reference stops and leaves the same
// partial tree.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"
#include "tree_dev.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned long long kNone = ~0ull;   // key of a particle that descends no further
constexpr unsigned kAlign = 256;
constexpr unsigned kMaxLevels = 17;           // depths 0..16: tree_depth_max is at most 16
constexpr unsigned kLevelBatch = 4;           // levels queued per host synchronisation

// Device build state st[]: the level's slot count, level-major id of its
// first slot, splits of the shallower levels, particles still descending (a
// bound at depth 0), first archive entry, status, depth, and the particles
// placed in the walk order from the front (leaf members) and from the back
// (particles outside the bounds).
enum State { ST_M, ST_BASE, ST_EVOFF, ST_N, ST_AOFF, ST_STATUS, ST_DEPTH, ST_DONE, ST_TAIL, kStateLen };
enum Status { RUNNING = 0, DONE = 1, FALLBACK = 2 };
static_assert(kStateLen <= SF_TREE_PAR_STATE, "the host mirror holds the whole state");
// Level counters lm[]: splits, particles that descend, split at depth_max.
enum Level { LM_SPLITS, LM_NEXT, LM_DEEP, kLevelLen };

__device__ __forceinline__ void child_bounds(const float* bmin, const float* bmax, const float* mid, int o, float* out) {
	for (int a = 0; a < 3; ++a) {
		out[a] = ((o >> a) & 1) ? mid[a] : bmin[a];
		out[3 + a] = ((o >> a) & 1) ? bmax[a] : mid[a];
	}
}

__device__ __forceinline__ void clear_node(sf_node& n, unsigned parent, const float* b) {
	n.parent = parent;
	n.bodies = 0u;
	for (int k = 0; k < 8; ++k) n.branch[k] = 0u;
	for (int k = 0; k < 6; ++k) n.bounds[k] = b[k];
	for (int k = 0; k < 4; ++k) n.params[k] = 0.0f;
}

// The root record as the reference purges it, and the build state.
__global__ void k_par_state(unsigned N, const float* __restrict__ tb, sf_node* __restrict__ tnode,
                            unsigned* __restrict__ st, unsigned* __restrict__ lm) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	float b[6];
	for (int k = 0; k < 6; ++k) b[k] = tb[k];
	clear_node(tnode[0], 0u, b);
	for (int k = 0; k < kStateLen; ++k) st[k] = 0u;
	st[ST_M] = 1u;
	st[ST_N] = N;
	st[ST_STATUS] = RUNNING;
	for (int k = 0; k < kLevelLen; ++k) lm[k] = 0u;
}

// key = 0 for particles inside the tree bounds (the test of src/tree.c:147-152,
// so NaN coordinates are inside), kNone otherwise and for the padding;
// particles outside take the back of the walk order.
__global__ __launch_bounds__(kBlock) void k_par_init(const float* __restrict__ pos, unsigned N, unsigned n_sort,
                                                     const float* __restrict__ tb, unsigned long long* __restrict__ key,
                                                     unsigned* __restrict__ val, unsigned* __restrict__ st,
                                                     unsigned* __restrict__ order) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_sort) return;
	unsigned long long k = kNone;
	if (i < N) {
		const float px = pos[3u * i], py = pos[3u * i + 1u], pz = pos[3u * i + 2u];
		if (!(tb[0] > px || px > tb[3] || tb[1] > py || py > tb[4] || tb[2] > pz || pz > tb[5])) k = 0ull;
		else order[N - 1u - atomicAdd(&st[ST_TAIL], 1u)] = i;
	}
	key[i] = k;
	val[i] = i;
}

// seg[s] = (0, 0) for the level's slots
__global__ __launch_bounds__(kBlock) void k_seg_clear(const unsigned* __restrict__ st, uint2* __restrict__ seg) {
	const unsigned s = blockIdx.x * kBlock + threadIdx.x;
	if (st[ST_STATUS] != RUNNING || s >= st[ST_M]) return;
	seg[s] = make_uint2(0u, 0u);
}

// [first, end) of each slot's run of sorted keys, and the member products
// (m x, m y, m z, m) of the level in sorted order, one component per archive
// plane of a_cap floats, from entry st[ST_AOFF].
__global__ __launch_bounds__(kBlock) void k_seg_mark(const unsigned* __restrict__ st, const float* __restrict__ pos,
                                                     const float* __restrict__ mas,
                                                     const unsigned long long* __restrict__ key,
                                                     const unsigned* __restrict__ val, size_t a_cap,
                                                     uint2* __restrict__ seg, float* __restrict__ arch) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	const unsigned n = st[ST_N];
	if (st[ST_STATUS] != RUNNING || p >= n) return;
	const unsigned long long k = key[p];
	if (k == kNone) return;
	const unsigned s = (unsigned)k;
	if (p == 0u || key[p - 1u] != k) seg[s].x = p;
	if (p + 1u == n || key[p + 1u] != k) seg[s].y = p + 1u;
	const size_t q = (size_t)st[ST_AOFF] + p;
	const unsigned i = val[p];
	const float pm = mas[i];
	arch[q] = pm * pos[3u * i];
	arch[a_cap + q] = pm * pos[3u * i + 1u];
	arch[2u * a_cap + q] = pm * pos[3u * i + 2u];
	arch[3u * a_cap + q] = pm;
}

// Body counts, leaf lists and archive ranges (trange) of the level's nodes;
// flag[s] = 1 for a splitting slot.  A leaf's members also take the next
// places of the walk order, so that a wavefront of the walk holds particles
// of neighbouring leaves (any order gives the same per-particle results).
__global__ __launch_bounds__(kBlock) void k_slot_classify(unsigned* __restrict__ st,
                                                          const unsigned* __restrict__ val, const uint2* __restrict__ seg,
                                                          unsigned depth_max, sf_node* __restrict__ tnode,
                                                          uint2* __restrict__ trange, unsigned* __restrict__ flag,
                                                          unsigned* __restrict__ lm, unsigned* __restrict__ order) {
	const unsigned s = blockIdx.x * kBlock + threadIdx.x;
	if (st[ST_STATUS] != RUNNING || s >= st[ST_M]) return;
	const unsigned t = st[ST_BASE] + s, off = st[ST_AOFF];
	const uint2 r = seg[s];
	const unsigned c = r.y - r.x;
	sf_node& nd = tnode[t];
	nd.bodies = c;
	trange[t] = make_uint2(off + r.x, off + r.y);
	if (c >= 9u) {
		flag[s] = 1u;
		atomicAdd(&lm[LM_SPLITS], 1u);
		atomicAdd(&lm[LM_NEXT], c);
		if (st[ST_DEPTH] >= depth_max) lm[LM_DEEP] = 1u;
	} else {
		flag[s] = 0u;
		const unsigned w = c ? atomicAdd(&st[ST_DONE], c) : 0u;
		for (unsigned k = 0; k < c; ++k) {
			const unsigned i = val[r.x + k];
			nd.branch[k] = i;
			order[w + k] = i;
		}
	}
}

// Exclusive scan of the level's split flags in one block: each thread scans
// a contiguous chunk, then the chunk totals.
__global__ __launch_bounds__(1024) void k_flag_scan(const unsigned* __restrict__ st, unsigned* __restrict__ flag) {
	__shared__ unsigned tot[1024];
	if (st[ST_STATUS] != RUNNING) return;
	const unsigned n = st[ST_M];
	const unsigned chunk = (n + 1023u) / 1024u;
	const unsigned b = threadIdx.x * chunk, e = min(b + chunk, n);
	unsigned s = 0;
	for (unsigned k = b; k < e; ++k) s += flag[k];
	tot[threadIdx.x] = s;
	__syncthreads();
	for (unsigned off = 1; off < 1024u; off <<= 1) {
		const unsigned v = threadIdx.x >= off ? tot[threadIdx.x - off] : 0u;
		__syncthreads();
		tot[threadIdx.x] += v;
		__syncthreads();
	}
	unsigned run = threadIdx.x ? tot[threadIdx.x - 1u] : 0u;
	for (unsigned k = b; k < e; ++k) {
		const unsigned v = flag[k];
		flag[k] = run;
		run += v;
	}
}

// Next key of each member of a splitting slot: 8 * (split number within the
// level) + octant, by '<' for the slot's first eight members and '<=' after
// them; members of leaves get kNone.  The slot's rank-8 member (i9) also
// creates the children at level-major ids 1 + 8 g + octant (g = splits
// before this one, in level-major order), sets the parent's branch and
// records the split event (i9 << 8 | depth, g).  Writes stay inside the node
// and event capacities; a build that reaches them falls back
// (k_level_advance).
__global__ __launch_bounds__(kBlock) void k_route(const unsigned* __restrict__ st, const float* __restrict__ pos,
                                                  const unsigned* __restrict__ val, const uint2* __restrict__ seg,
                                                  const unsigned* __restrict__ split_local, unsigned node_cap,
                                                  unsigned ev_cap, sf_node* __restrict__ tnode,
                                                  unsigned long long* __restrict__ key,
                                                  unsigned long long* __restrict__ ev_key,
                                                  unsigned* __restrict__ ev_val) {
	const unsigned p = blockIdx.x * kBlock + threadIdx.x;
	if (st[ST_STATUS] != RUNNING || p >= st[ST_N]) return;
	const unsigned long long k = key[p];
	if (k == kNone) return;
	const unsigned s = (unsigned)k;
	const uint2 r = seg[s];
	if (r.y - r.x < 9u) {
		key[p] = kNone;
		return;
	}
	const unsigned i = val[p];
	const unsigned t = st[ST_BASE] + s;
	sf_node& nd = tnode[t];
	float bmin[3], bmax[3], mid[3];
	for (int a = 0; a < 3; ++a) {
		bmin[a] = nd.bounds[a];
		bmax[a] = nd.bounds[3 + a];
		mid[a] = 0.5f * (bmin[a] + bmax[a]);
	}
	const float x[3] = {pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u]};
	const unsigned rank = p - r.x;
	int oct = 0;
	if (rank < 8u) {
		for (int a = 0; a < 3; ++a) oct += mid[a] < x[a] ? (1 << a) : 0;
	} else {
		for (int a = 0; a < 3; ++a) oct += mid[a] <= x[a] ? (1 << a) : 0;
	}
	key[p] = 8ull * split_local[s] + (unsigned)oct;
	if (rank != 8u) return;
	const unsigned g = st[ST_EVOFF] + split_local[s];
	if (g >= ev_cap || 9ull + 8ull * g > (unsigned long long)node_cap) return;
	const unsigned child0 = 1u + 8u * g;
	for (int o = 0; o < 8; ++o) {
		float b[6];
		child_bounds(bmin, bmax, mid, o, b);
		clear_node(tnode[child0 + (unsigned)o], t, b);
		nd.branch[o] = child0 + (unsigned)o;
	}
	ev_key[g] = ((unsigned long long)i << 8) | st[ST_DEPTH];
	ev_val[g] = g;
}

// End of a level: stop (no split), fall back (a split at depth_max, or more
// than max_nodes nodes), or step the state to the next level.
__global__ void k_level_advance(unsigned* __restrict__ st, unsigned* __restrict__ lm, unsigned max_nodes) {
	if (blockIdx.x != 0 || threadIdx.x != 0 || st[ST_STATUS] != RUNNING) return;
	const unsigned splits = lm[LM_SPLITS];
	if (lm[LM_DEEP] || 1ull + 8ull * ((unsigned long long)st[ST_EVOFF] + splits) >= (unsigned long long)max_nodes) {
		st[ST_STATUS] = FALLBACK;
	} else if (splits == 0u) {
		st[ST_STATUS] = DONE;
	} else {
		st[ST_AOFF] += st[ST_N];
		st[ST_BASE] += st[ST_M];
		st[ST_EVOFF] += splits;
		st[ST_M] = 8u * splits;
		st[ST_N] = lm[LM_NEXT];
		st[ST_DEPTH] += 1u;
	}
	for (int k = 0; k < kLevelLen; ++k) lm[k] = 0u;
}

// Moment sums params[c] = (((0 + q_0) + q_1) + ...) over each node's members
// in index order: the FP32 chain of the serial insertion (src/tree.c:205-208).
// Nodes of more than min_size members: one wavefront per (node, component),
// lanes loading 64 products at a time two chunks ahead and the chain
// reading them back lane by lane; waves stride over the (node, component)
// items.
constexpr unsigned kSumSmall = 256;   // larger nodes are summed by a wavefront

__global__ __launch_bounds__(64) void k_node_sums(const float* __restrict__ arch, size_t a_cap,
                                                  const uint2* __restrict__ trange, unsigned first, unsigned count,
                                                  unsigned min_size, sf_node* __restrict__ tnode) {
	const unsigned lane = threadIdx.x;
	for (unsigned item = blockIdx.x; item < 4u * count; item += gridDim.x) {
		const unsigned t = first + (item >> 2), comp = item & 3u;
		const uint2 r = trange[t];
		if (r.y - r.x <= min_size || r.y == r.x) continue;
		const float* __restrict__ a = arch + comp * a_cap;
		float v0 = r.x + lane < r.y ? a[r.x + lane] : 0.0f;
		float v1 = r.x + 64u + lane < r.y ? a[r.x + 64u + lane] : 0.0f;
		float sum = 0.0f;
		for (unsigned c = r.x; c < r.y; c += 64u) {
			const float cur = v0;
			v0 = v1;
			v1 = c + 128u + lane < r.y ? a[c + 128u + lane] : 0.0f;
			const unsigned m = r.y - c;
			if (m >= 64u) {
#pragma unroll
				for (int l = 0; l < 64; ++l)
					sum = sum + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur), l));
			} else {
				for (unsigned l = 0; l < m; ++l)
					sum = sum + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur), (int)l));
			}
		}
		if (lane == 0u) tnode[t].params[comp] = sum;
	}
}

// The same sums for nodes of 1..kSumSmall members, one lane per
// (node, component).
__global__ __launch_bounds__(kBlock) void k_node_sums_small(const float* __restrict__ arch, size_t a_cap,
                                                            const uint2* __restrict__ trange, unsigned first,
                                                            unsigned count, sf_node* __restrict__ tnode) {
	const unsigned item = blockIdx.x * kBlock + threadIdx.x;
	if (item >= 4u * count) return;
	const unsigned t = first + (item >> 2), comp = item & 3u;
	const uint2 r = trange[t];
	if (r.y == r.x || r.y - r.x > kSumSmall) return;
	const float* __restrict__ a = arch + comp * a_cap;
	float sum = 0.0f;
	for (unsigned i = r.x; i < r.y; ++i) sum = sum + a[i];
	tnode[t].params[comp] = sum;
}

__global__ __launch_bounds__(kBlock) void k_fill_none(unsigned long long* __restrict__ key, unsigned b, unsigned e) {
	const unsigned i = b + blockIdx.x * kBlock + threadIdx.x;
	if (i < e) key[i] = kNone;
}

// perm[g] = position of split g in (i9, depth) order
__global__ __launch_bounds__(kBlock) void k_ev_perm(const unsigned* __restrict__ ev_sorted, unsigned S,
                                                    unsigned* __restrict__ perm) {
	const unsigned r = blockIdx.x * kBlock + threadIdx.x;
	if (r < S) perm[ev_sorted[r]] = r;
}

__device__ __forceinline__ unsigned final_id(unsigned t, const unsigned* __restrict__ perm) {
	return t == 0u ? 0u : 1u + 8u * perm[(t - 1u) >> 3] + ((t - 1u) & 7u);
}

// node[final(t)] = tnode[t] with parent and child ids renumbered
__global__ __launch_bounds__(kBlock) void k_final_nodes(const sf_node* __restrict__ tnode, unsigned num_nodes,
                                                        const unsigned* __restrict__ perm, sf_node* __restrict__ node,
                                                        unsigned* __restrict__ meta) {
	const unsigned t = blockIdx.x * kBlock + threadIdx.x;
	if (t == 0u) {
		meta[0] = num_nodes;
		meta[1] = 0u;
	}
	if (t >= num_nodes) return;
	sf_node rec = tnode[t];
	rec.parent = final_id(rec.parent, perm);
	if (rec.bodies > 8u)
		for (int o = 0; o < 8; ++o) rec.branch[o] = final_id(rec.branch[o], perm);
	node[final_id(t, perm)] = rec;
}

// Centre of mass, as k_tree_com_ref (src/tree.c:302-310)
__global__ __launch_bounds__(kBlock) void k_par_com(sf_node* __restrict__ node, unsigned num_nodes) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= num_nodes) return;
	const float mass = node[i].params[3];
	const float inv = 0.0f < mass ? 1.0f / mass : 1.0f;
	node[i].params[0] = node[i].params[0] * inv;
	node[i].params[1] = node[i].params[1] * inv;
	node[i].params[2] = node[i].params[2] * inv;
}

inline unsigned blocks(unsigned n) { return (n + kBlock - 1u) / kBlock; }
inline unsigned round_up(unsigned n, unsigned m) { return (n + m - 1u) / m * m; }
inline size_t align_up(size_t n) { return (n + kAlign - 1u) / kAlign * kAlign; }

// 8-bit radix passes for keys up to kmax, kept below kNone's low bytes
inline unsigned passes_for(unsigned long long kmax) {
	unsigned p = 1u;
	while (p < 8u && kmax >= (1ull << (8u * p)) - 1ull) ++p;
	return p;
}

// Scratch planes, each 256-byte aligned.
enum Plane { KEY0, KEY1, VAL0, VAL1, EVK0, EVK1, EVV0, EVV1, HIST, SEG, FLAG, TNODE, TRANGE, PERM, STATE, LMETA, ARCH,
             kPlanes };

struct Layout {
	unsigned n_sort, ev_cap, ev_sort, slot_cap;
	size_t a_cap;   // floats per product plane
	size_t off[kPlanes];
	size_t bytes;
};

Layout layout(unsigned N, unsigned max_nodes) {
	Layout L;
	const unsigned tile = sfd_radix_tile();
	const size_t nodes = max_nodes ? max_nodes : 1u;
	L.n_sort = round_up(N ? N : 1u, tile);
	L.ev_cap = max_nodes / 8u + 1u;
	L.ev_sort = round_up(L.ev_cap, tile);
	L.slot_cap = 8u * L.ev_cap + 8u;
	L.a_cap = (size_t)kMaxLevels * (N ? N : 1u);
	const unsigned hist_n = sfd_radix_hist_entries(L.n_sort > L.ev_sort ? L.n_sort : L.ev_sort);
	size_t sz[kPlanes];
	sz[KEY0] = sz[KEY1] = 8u * (size_t)L.n_sort;
	sz[VAL0] = sz[VAL1] = 4u * (size_t)L.n_sort;
	sz[EVK0] = sz[EVK1] = 8u * (size_t)L.ev_sort;
	sz[EVV0] = sz[EVV1] = 4u * (size_t)L.ev_sort;
	sz[HIST] = 4u * (size_t)hist_n;
	sz[SEG] = 8u * (size_t)L.slot_cap;
	sz[FLAG] = 4u * (size_t)L.slot_cap;
	sz[TNODE] = sizeof(sf_node) * nodes;
	sz[TRANGE] = 8u * nodes;
	sz[PERM] = 4u * (size_t)L.ev_cap;
	sz[STATE] = 4u * kStateLen;
	sz[LMETA] = 4u * kLevelLen;
	sz[ARCH] = 4u * 4u * L.a_cap;
	size_t o = 0;
	for (int k = 0; k < kPlanes; ++k) {
		L.off[k] = o;
		o += align_up(sz[k]);
	}
	L.bytes = o;
	return L;
}

template <class T>
T* plane(void* scratch, const Layout& L, Plane k) {
	return reinterpret_cast<T*>(static_cast<char*>(scratch) + L.off[k]);
}

// The level-synchronous build applies when the tree can hold N bodies (at
// most 8 per leaf, so N <= 8 max_nodes) and archive offsets fit 32 bits.
bool par_applies(unsigned N, unsigned max_nodes) {
	return max_nodes >= 2u && (unsigned long long)N <= 8ull * max_nodes &&
	       (unsigned long long)kMaxLevels * N < (1ull << 32);
}

// Moment sums of nodes [first, first + count): lanes for the small ones,
// wavefronts for the others.
void launch_sums(const float* arch, size_t a_cap, const uint2* trange, unsigned first, unsigned count, sf_node* tnode,
                 hipStream_t s) {
	const unsigned items = 4u * count;
	hipLaunchKernelGGL(k_node_sums_small, dim3(blocks(items)), dim3(kBlock), 0, s, arch, a_cap, trange, first, count,
	                   tnode);
	hipLaunchKernelGGL(k_node_sums, dim3(items < 4096u ? items : 4096u), dim3(64), 0, s, arch, a_cap, trange, first,
	                   count, kSumSmall, tnode);
}

int build_par(const sf_tree_par_ctx* x, const float* pos, const float* mas, unsigned N, const float* bounds,
              unsigned max_nodes, int depth_max, sf_node* node, unsigned* meta, unsigned* order, hipStream_t s) {
	if (!par_applies(N, max_nodes) || depth_max < 0 || depth_max >= (int)kMaxLevels) return 1;
	const Layout L = layout(N, max_nodes);
	void* scr = x->scratch;
	unsigned long long* key[2] = {plane<unsigned long long>(scr, L, KEY0), plane<unsigned long long>(scr, L, KEY1)};
	unsigned* val[2] = {plane<unsigned>(scr, L, VAL0), plane<unsigned>(scr, L, VAL1)};
	unsigned long long* ev_key[2] = {plane<unsigned long long>(scr, L, EVK0), plane<unsigned long long>(scr, L, EVK1)};
	unsigned* ev_val[2] = {plane<unsigned>(scr, L, EVV0), plane<unsigned>(scr, L, EVV1)};
	unsigned* hist = plane<unsigned>(scr, L, HIST);
	uint2* seg = plane<uint2>(scr, L, SEG);
	unsigned* flag = plane<unsigned>(scr, L, FLAG);
	sf_node* tnode = plane<sf_node>(scr, L, TNODE);
	uint2* trange = plane<uint2>(scr, L, TRANGE);
	unsigned* perm = plane<unsigned>(scr, L, PERM);
	unsigned* st = plane<unsigned>(scr, L, STATE);
	unsigned* lm = plane<unsigned>(scr, L, LMETA);
	float* arch = plane<float>(scr, L, ARCH);

	hipLaunchKernelGGL(k_par_state, dim3(1), dim3(64), 0, s, N, bounds, tnode, st, lm);
	hipLaunchKernelGGL(k_par_init, dim3(blocks(L.n_sort)), dim3(kBlock), 0, s, pos, N, L.n_sort, bounds, key[0],
	                   val[0], st, order);
	// inside particles first, in index order
	if (sfd_radix_sort_passes(key[0], val[0], key[1], val[1], L.n_sort, hist, 1u, s)) return -1;
	int cur = 1;
	// launch bounds: at most N particles descend, a level has at most
	// 8 (N / 9) + 1 slots, and next keys stay below 8 (N / 9)
	const unsigned n_grid = blocks(N ? N : 1u);
	const unsigned m_grid = blocks(8u * (N / 9u) + 1u);
	const unsigned passes = passes_for(8ull * (N / 9u));
	const unsigned n_pad = round_up(N ? N : 1u, sfd_radix_tile());
	unsigned levels = 0;
	unsigned summed = 1u;   // nodes [1, summed) have their sums queued on the side stream
	while (levels < kMaxLevels) {
		const unsigned batch = kMaxLevels - levels < kLevelBatch ? kMaxLevels - levels : kLevelBatch;
		for (unsigned b = 0; b < batch; ++b, ++levels) {
			hipLaunchKernelGGL(k_seg_clear, dim3(m_grid), dim3(kBlock), 0, s, st, seg);
			hipLaunchKernelGGL(k_seg_mark, dim3(n_grid), dim3(kBlock), 0, s, st, pos, mas, key[cur], val[cur], L.a_cap,
			                   seg, arch);
			hipLaunchKernelGGL(k_slot_classify, dim3(m_grid), dim3(kBlock), 0, s, st, val[cur], seg,
			                   (unsigned)depth_max, tnode, trange, flag, lm, order);
			if (levels == 0) {
				// the root's sums, the longest chain, run beside the deeper levels
				if (sf_hip_check(hipEventRecord(x->ev[0], s), "sfd_tree_build_ref_par", "event") ||
				    sf_hip_check(hipStreamWaitEvent(x->side, x->ev[0], 0), "sfd_tree_build_ref_par", "wait"))
					return -1;
				hipLaunchKernelGGL(k_node_sums, dim3(4), dim3(64), 0, x->side, arch, L.a_cap, trange, 0u, 1u, 0u, tnode);
				if (sf_hip_check(hipEventRecord(x->ev[1], x->side), "sfd_tree_build_ref_par", "event")) return -1;
			}
			hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(1024), 0, s, st, flag);
			hipLaunchKernelGGL(k_route, dim3(n_grid), dim3(kBlock), 0, s, st, pos, val[cur], seg, flag, max_nodes,
			                   L.ev_cap, tnode, key[cur], ev_key[0], ev_val[0]);
			if (sfd_radix_sort_passes(key[cur], val[cur], key[cur ^ 1], val[cur ^ 1], n_pad, hist, passes, s))
				return -1;
			if (passes & 1u) cur ^= 1;
			hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(64), 0, s, st, lm, max_nodes);
		}
		if (sf_hip_check(hipMemcpyAsync(x->h_state, st, kStateLen * sizeof(unsigned), hipMemcpyDeviceToHost, s),
		                 "sfd_tree_build_ref_par", "state copy") ||
		    sf_hip_check(hipStreamSynchronize(s), "sfd_tree_build_ref_par", "levels"))
			return -1;
		if (x->h_state[ST_STATUS] != RUNNING) break;
		// nodes of the finished levels (all before the current level's first
		// slot) are complete: their sums run beside the next levels
		const unsigned upto = x->h_state[ST_BASE];
		if (upto > summed) {
			launch_sums(arch, L.a_cap, trange, summed, upto - summed, tnode, x->side);
			if (sf_hip_check(hipEventRecord(x->ev[1], x->side), "sfd_tree_build_ref_par", "event")) return -1;
			summed = upto;
		}
	}
	if (x->h_state[ST_STATUS] != DONE) {
		// sums on the side stream may still read the archive: finish them first
		if (sf_hip_check(hipStreamSynchronize(x->side), "sfd_tree_build_ref_par", "side")) return -1;
		return 1;
	}

	const unsigned S = x->h_state[ST_EVOFF];
	const unsigned num_nodes = 1u + 8u * S;
	if (num_nodes > summed) launch_sums(arch, L.a_cap, trange, summed, num_nodes - summed, tnode, s);
	if (S) {
		const unsigned ev_pad = round_up(S, sfd_radix_tile());
		if (ev_pad > S)
			hipLaunchKernelGGL(k_fill_none, dim3(blocks(ev_pad - S)), dim3(kBlock), 0, s, ev_key[0], S, ev_pad);
		const unsigned long long kmax = ((unsigned long long)(N ? N - 1u : 0u) << 8) | 0xFFull;
		const unsigned ev_passes = passes_for(kmax);
		if (sfd_radix_sort_passes(ev_key[0], ev_val[0], ev_key[1], ev_val[1], ev_pad, hist, ev_passes, s)) return -1;
		hipLaunchKernelGGL(k_ev_perm, dim3(blocks(S)), dim3(kBlock), 0, s, ev_val[ev_passes & 1u], S, perm);
	}
	if (sf_hip_check(hipStreamWaitEvent(s, x->ev[1], 0), "sfd_tree_build_ref_par", "wait")) return -1;
	hipLaunchKernelGGL(k_final_nodes, dim3(blocks(num_nodes)), dim3(kBlock), 0, s, tnode, num_nodes, perm, node, meta);
	hipLaunchKernelGGL(k_par_com, dim3(blocks(num_nodes)), dim3(kBlock), 0, s, node, num_nodes);
	return sf_hip_check(hipGetLastError(), "sfd_tree_build_ref_par", "launch") ? -1 : 0;
}

}  // namespace

extern "C" size_t sfd_tree_par_bytes(unsigned N, unsigned max_nodes) {
	return par_applies(N, max_nodes) ? layout(N, max_nodes).bytes : 0u;
}

extern "C" int sfd_tree_build_ref_par(const sf_tree_par_ctx* x, const float* pos, const float* mas, unsigned N,
                                      const float* bounds, unsigned max_nodes, int depth_max, sf_node* node,
                                      unsigned* meta, unsigned* order, hipStream_t s) {
	const int tok = sf_prof_begin("tree_build_ref_par", s);
	const int rc = build_par(x, pos, mas, N, bounds, max_nodes, depth_max, node, meta, order, s);
	sf_prof_end(tok, s);
	return rc;
}
