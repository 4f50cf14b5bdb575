// tree_ref.hip -- the reference's octree (src/tree.c:74-318) and its walk
// (src/gravity.c:245-663), reproduced bit for bit on the device.
//
// The reference inserts particles one at a time in index order; node ids,
// the per-node FP32 mass-moment sums and the leaf particle order all depend
// on that order, so the build is one lane walking the particles in order
// (SF_TREE_LBVH is the parallel build).  The walk is one lane per particle
// with the reference's parent-pointer walk, opening test and leaf-monopole
// self-removal, in the reference's FP32 operation order (the file is
// compiled with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "../sf_internal.h"
#include "tree_dev.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr int kDepthCap = 16;   // TREE_DEPTH_MAX upper limit (tree_path size)

// Radii r_i = sqrt((x^2 + y^2) + z^2) in FP64 (src/simulation.c:256-262),
// all particles at once.
__global__ __launch_bounds__(kBlock) void k_radius_f64(const float* __restrict__ pos, unsigned N,
                                                       double* __restrict__ r1) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= N) return;
	const double x = pos[3u * i], y = pos[3u * i + 1u], z = pos[3u * i + 2u];
	r1[i] = sqrt((x * x + y * y) + z * z);
}

__device__ __forceinline__ void kahan_step(double& avg, double& com, double r) {
	const double yk = r - com;
	const double t = avg + yk;
	const double zk = t - avg;
	com = zk - yk;
	avg = t;
}

// Mean radius: the sequential FP64 Kahan sum of src/simulation.c:245-277
// (and src/visualization.c:276-305) over the radii, in one wavefront; the
// radii arrive through scalar loads, 16 ahead of the chain.
constexpr unsigned kKahanAhead = 16;

__device__ double kahan_mean(const double* __restrict__ r1, unsigned N) {
	double avg = 0.0, com = 0.0;
	unsigned i = 0;
	if (N >= kKahanAhead) {
		double nx[kKahanAhead];
#pragma unroll
		for (unsigned u = 0; u < kKahanAhead; ++u) nx[u] = r1[u];
		for (; i + 2u * kKahanAhead <= N; i += kKahanAhead) {
			double cur[kKahanAhead];
#pragma unroll
			for (unsigned u = 0; u < kKahanAhead; ++u) cur[u] = nx[u];
#pragma unroll
			for (unsigned u = 0; u < kKahanAhead; ++u) nx[u] = r1[i + kKahanAhead + u];
#pragma unroll
			for (unsigned u = 0; u < kKahanAhead; ++u) kahan_step(avg, com, cur[u]);
		}
#pragma unroll
		for (unsigned u = 0; u < kKahanAhead; ++u) kahan_step(avg, com, nx[u]);
		i += kKahanAhead;
	}
	for (; i < N; ++i) kahan_step(avg, com, r1[i]);
	return avg * (1.0 / (double)N);
}

// The jittered cube bounds (src/simulation.c:279-301) around the mean
// radius; jitter[i] is 2 * (INV_PCG32_MAX * s[i]) - 1 from pcg4d of the step
// number.
__global__ __launch_bounds__(64) void k_tree_bounds_ref(const double* __restrict__ r1, unsigned N, double j0, double j1,
                                                        double j2, float* __restrict__ bounds) {
	const double avg = kahan_mean(r1, N);
	if (threadIdx.x != 0) return;
	const double jit[3] = {j0, j1, j2};
	for (int a = 0; a < 3; ++a) {
		bounds[a] = (float)(-2.000 * avg + 0.5 * jit[a] * avg);
		bounds[3 + a] = (float)(2.000 * avg + 0.5 * jit[a] * avg);
	}
}

__global__ __launch_bounds__(64) void k_mean_radius(const double* __restrict__ r1, unsigned N,
                                                    double* __restrict__ avg) {
	const double a = kahan_mean(r1, N);
	if (threadIdx.x == 0) avg[0] = a;
}

// Serial insertion (src/tree.c:120-296).  meta[0] = num_nodes,
// meta[1] = status (0 ok, 1 out of nodes, 2 too deep).
__global__ void k_tree_build_ref(const float* __restrict__ pos, const float* __restrict__ mas, unsigned N,
                                 const float* __restrict__ tbounds, unsigned max_nodes, int depth_max,
                                 sf_node* __restrict__ node, unsigned* __restrict__ meta) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	unsigned num_nodes = 1u;
	float tb[6];
	for (int k = 0; k < 6; ++k) tb[k] = tbounds[k];
	// purge of the root (nodes past num_nodes are never read)
	{
		sf_node& r = node[0];
		r.parent = 0u;
		r.bodies = 0u;
		for (int k = 0; k < 8; ++k) r.branch[k] = 0u;
		for (int k = 0; k < 6; ++k) r.bounds[k] = tb[k];
		for (int k = 0; k < 4; ++k) r.params[k] = 0.0f;
	}
	unsigned status = 0u;
	for (unsigned idx = 0; idx < N && status == 0u; ++idx) {
		const float px = pos[3u * idx], py = pos[3u * idx + 1u], pz = pos[3u * idx + 2u];
		const float pm = mas[idx];
		float bmin[3] = {tb[0], tb[1], tb[2]};
		float bmax[3] = {tb[3], tb[4], tb[5]};
		const float p[3] = {px, py, pz};
		if (bmin[0] > px || px > bmax[0] || bmin[1] > py || py > bmax[1] || bmin[2] > pz || pz > bmax[2]) continue;
		unsigned cur = 0u;
		for (int depth = 0; depth <= depth_max; ++depth) {
			sf_node& cn = node[cur];
			const unsigned cur_bodies = cn.bodies;
			cn.bodies = cur_bodies + 1u;
			float mid[3];
			for (int a = 0; a < 3; ++a) mid[a] = 0.5f * (bmin[a] + bmax[a]);
			cn.params[0] = cn.params[0] + pm * px;
			cn.params[1] = cn.params[1] + pm * py;
			cn.params[2] = cn.params[2] + pm * pz;
			cn.params[3] = cn.params[3] + pm;
			if (cur_bodies < 8u) {
				cn.branch[cur_bodies] = idx;
				break;
			}
			if (cur_bodies == 8u) {
				if (max_nodes <= num_nodes + 8u) {
					status = 1u;
					break;
				}
				if (depth_max <= depth) {
					status = 2u;
					break;
				}
				for (int o = 0; o < 8; ++o) {
					sf_node& nn = node[num_nodes + (unsigned)o];
					nn.parent = cur;
					nn.bodies = 0u;
					for (int k = 0; k < 8; ++k) nn.branch[k] = 0u;
					for (int a = 0; a < 3; ++a) {
						nn.bounds[a] = ((o >> a) & 1) ? mid[a] : bmin[a];
						nn.bounds[3 + a] = ((o >> a) & 1) ? bmax[a] : mid[a];
					}
					for (int k = 0; k < 4; ++k) nn.params[k] = 0.0f;
				}
				// redistribution uses '<' (src/tree.c:244-247)
				for (int k = 0; k < 8; ++k) {
					const unsigned ci = cn.branch[k];
					const float cx = pos[3u * ci], cy = pos[3u * ci + 1u], cz = pos[3u * ci + 2u];
					const float cm = mas[ci];
					const int oc = (mid[0] < cx ? 1 : 0) + (mid[1] < cy ? 2 : 0) + (mid[2] < cz ? 4 : 0);
					sf_node& ch = node[num_nodes + (unsigned)oc];
					const unsigned cb = ch.bodies;
					ch.bodies = cb + 1u;
					ch.branch[cb] = ci;
					ch.params[0] = ch.params[0] + cm * cx;
					ch.params[1] = ch.params[1] + cm * cy;
					ch.params[2] = ch.params[2] + cm * cz;
					ch.params[3] = ch.params[3] + cm;
				}
				for (int o = 0; o < 8; ++o) cn.branch[o] = num_nodes + (unsigned)o;
				num_nodes += 8u;
			}
			// insertion uses '<=' (src/tree.c:282-285)
			const int oct = (mid[0] <= p[0] ? 1 : 0) + (mid[1] <= p[1] ? 2 : 0) + (mid[2] <= p[2] ? 4 : 0);
			for (int a = 0; a < 3; ++a) {
				const bool hi = (oct >> a) & 1;
				bmin[a] = hi ? mid[a] : bmin[a];
				bmax[a] = hi ? bmax[a] : mid[a];
			}
			cur = cn.branch[oct];
		}
	}
	meta[0] = num_nodes;
	meta[1] = status;
}

// Centre of mass: moment sums divided by the mass (src/tree.c:302-310);
// skipped after a failed build, as the reference returns before it.
__global__ __launch_bounds__(kBlock) void k_tree_com_ref(sf_node* __restrict__ node, const unsigned* __restrict__ meta) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	if (meta[1] != 0u || i >= meta[0]) return;
	const float mass = node[i].params[3];
	const float inv = 0.0f < mass ? 1.0f / mass : 1.0f;
	node[i].params[0] = node[i].params[0] * inv;
	node[i].params[1] = node[i].params[1] * inv;
	node[i].params[2] = node[i].params[2] * inv;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
	const float y = x - c;
	const float t = s + y;
	const float z = t - s;
	c = z - y;
	s = t;
}

// The opening test of src/gravity.c:334-357 for a node of more than eight
// bodies: open iff theta <= (mean width)^2 / (distance to the centre)^2.
__device__ __forceinline__ bool opens(const float* bmin, const float* bmax, const float* ri, float theta) {
	const float third = (float)(1.0 / 3.0);
	float off[3];
	for (int a = 0; a < 3; ++a) off[a] = 0.5f * (bmin[a] + bmax[a]) - ri[a];
	const float d2 = (off[0] * off[0] + off[1] * off[1]) + off[2] * off[2];
	const float w = third * (((bmax[0] - bmin[0]) + (bmax[1] - bmin[1])) + (bmax[2] - bmin[2]));
	const float ratio = 0.0f < d2 ? (w * w) / d2 : 0.0f;
	return theta <= ratio;
}

// The node as one monopole, less particle i when the node's box holds it
// (src/gravity.c:380-640), added into the (Kahan) sums.
template <bool KAHAN_ACC, bool KAHAN_POT>
__device__ __forceinline__ void add_monopole(const sf_node* __restrict__ nd, unsigned bodies, const float* bmin,
                                             const float* bmax, const float* ri, float mi, float eps2, float* sum,
                                             float* com) {
	float rj[3] = {nd->params[0], nd->params[1], nd->params[2]};
	float mj = nd->params[3];
	if (bmin[0] <= ri[0] && ri[0] <= bmax[0] && bmin[1] <= ri[1] && ri[1] <= bmax[1] && bmin[2] <= ri[2] &&
	    ri[2] <= bmax[2] && 1u < bodies) {
		for (int a = 0; a < 3; ++a) rj[a] = rj[a] * mj;
		for (int a = 0; a < 3; ++a) rj[a] = rj[a] - mi * ri[a];
		mj = mj - mi;
		for (int a = 0; a < 3; ++a) rj[a] = rj[a] * (0.0f < mj ? 1.0f / mj : 1.0f);
	}
	const float rx = ri[0] - rj[0], ry = ri[1] - rj[1], rz = ri[2] - rj[2];
	const float r2 = (rx * rx + ry * ry) + rz * rz;
	const float inv_r2 = 1.0f / (r2 + eps2);
	const float inv_r1 = 1.0f / sqrtf(r2 + eps2);
	const float pj = mj * inv_r1;
	const float f[3] = {pj * inv_r2 * rx, pj * inv_r2 * ry, pj * inv_r2 * rz};
	for (int a = 0; a < 3; ++a) {
		if (KAHAN_ACC) kahan_add(sum[a], com[a], f[a]);
		else sum[a] += f[a];
	}
	if (KAHAN_POT) kahan_add(sum[3], com[3], pj);
	else sum[3] += pj;
}

template <bool PHYS>
__device__ __forceinline__ void store_walk(unsigned idx, float mi, float negG, const float* sum,
                                           float* __restrict__ acc, float* __restrict__ pot) {
	const float f = PHYS ? negG : negG * mi;
	acc[3u * idx + 0u] = f * sum[0];
	acc[3u * idx + 1u] = f * sum[1];
	acc[3u * idx + 2u] = f * sum[2];
	pot[idx] = negG * sum[3];
}

// Per-particle walk of src/gravity.c:273-660: the reference's parent-pointer
// walk with its path stack and iteration cap.
template <bool KAHAN_ACC, bool KAHAN_POT, bool PHYS>
__global__ __launch_bounds__(kBlock) void k_tree_walk_ref(const float* __restrict__ pos, const float* __restrict__ mas,
                                                          unsigned N, const unsigned* __restrict__ order,
                                                          const sf_node* __restrict__ node,
                                                          unsigned node_cap, unsigned iter_max, float eps2,
                                                          float theta, float negG,
                                                          float* __restrict__ acc, float* __restrict__ pot) {
	const unsigned lane_id = blockIdx.x * kBlock + threadIdx.x;
	if (lane_id >= N) return;
	const unsigned idx = order ? order[lane_id] : lane_id;
	float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f}, com[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	const float ri[3] = {pos[3u * idx], pos[3u * idx + 1u], pos[3u * idx + 2u]};
	const float mi = mas[idx];
	unsigned next_node = 0u;
	int cur_depth = 0;
	int path[kDepthCap];
	for (int k = 0; k < kDepthCap; ++k) path[k] = -1;

	for (unsigned it = 0; it < iter_max; ++it) {
		if (cur_depth < 0) break;
		// Only a failed build (a node counted past 8 bodies but never split)
		// can lead outside the allocated nodes or below the path stack; the
		// reference reads out of bounds there, this walk stops.
		if (next_node >= node_cap) break;
		const sf_node* nd = &node[next_node];
		const unsigned bodies = nd->bodies;
		const unsigned parent = nd->parent;
		const float bmin[3] = {nd->bounds[0], nd->bounds[1], nd->bounds[2]};
		const float bmax[3] = {nd->bounds[3], nd->bounds[4], nd->bounds[5]};
		if (bodies == 0u) {
			next_node = parent;
			cur_depth -= 1;
			continue;
		}
		if (bodies > 8u) {
			if (cur_depth >= kDepthCap) break;   // only after a failed build
			const int octant = path[cur_depth];
			if (octant >= 7) {
				next_node = parent;
				path[cur_depth] = -1;
				cur_depth -= 1;
				continue;
			}
			if (octant < 0) {
				if (opens(bmin, bmax, ri, theta)) {
					next_node = nd->branch[0];
					path[cur_depth] = 0;
					cur_depth += 1;
					continue;
				}
			} else {
				next_node = nd->branch[octant + 1];
				path[cur_depth] = octant + 1;
				cur_depth += 1;
				continue;
			}
		}
		add_monopole<KAHAN_ACC, KAHAN_POT>(nd, bodies, bmin, bmax, ri, mi, eps2, sum, com);
		next_node = parent;
		cur_depth -= 1;
	}
	store_walk<PHYS>(idx, mi, negG, sum, acc, pot);
}

// The same walk over threaded links: down[n] is n's first non-empty child,
// next[n] the first non-empty node after n's subtree in the walk's order
// (children in octant order), kEnd past the last.  It visits the reference
// walk's non-empty nodes in its order without the returns to the parent;
// empty nodes add nothing there.  For a complete tree the reference walk
// takes 16 (opened nodes) + 1 <= 2 num_nodes - 1 iterations, below its cap,
// so the two walks add the same terms in the same order.
constexpr unsigned kEnd = 0xFFFFFFFFu;

template <bool KAHAN_ACC, bool KAHAN_POT, bool PHYS>
__global__ __launch_bounds__(kBlock) void k_tree_walk_thr(const float* __restrict__ pos, const float* __restrict__ mas,
                                                          unsigned N, const unsigned* __restrict__ order,
                                                          const sf_node* __restrict__ node,
                                                          const uint2* __restrict__ link, float eps2, float theta,
                                                          float negG, float* __restrict__ acc,
                                                          float* __restrict__ pot) {
	const unsigned lane_id = blockIdx.x * kBlock + threadIdx.x;
	if (lane_id >= N) return;
	const unsigned idx = order ? order[lane_id] : lane_id;
	float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f}, com[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	const float ri[3] = {pos[3u * idx], pos[3u * idx + 1u], pos[3u * idx + 2u]};
	const float mi = mas[idx];
	unsigned n = node[0].bodies ? 0u : kEnd;
	while (n != kEnd) {
		const sf_node* nd = &node[n];
		const uint2 ln = link[n];   // (down, next)
		const unsigned bodies = nd->bodies;
		const float bmin[3] = {nd->bounds[0], nd->bounds[1], nd->bounds[2]};
		const float bmax[3] = {nd->bounds[3], nd->bounds[4], nd->bounds[5]};
		if (bodies > 8u && opens(bmin, bmax, ri, theta)) {
			n = ln.x;
			continue;
		}
		add_monopole<KAHAN_ACC, KAHAN_POT>(nd, bodies, bmin, bmax, ri, mi, eps2, sum, com);
		n = ln.y;
	}
	store_walk<PHYS>(idx, mi, negG, sum, acc, pot);
}

// link[t] = (first non-empty child, first non-empty node after t's subtree)
// for the num_nodes nodes of a complete tree.  Walking up through last
// children takes at most the tree depth (16) steps.
__global__ __launch_bounds__(kBlock) void k_tree_links(const sf_node* __restrict__ node, const unsigned* __restrict__ meta,
                                                       uint2* __restrict__ link) {
	const unsigned t = blockIdx.x * kBlock + threadIdx.x;
	const unsigned num_nodes = meta[0];
	if (t >= num_nodes) return;
	unsigned down = kEnd;
	if (node[t].bodies > 8u)
		for (int k = 7; k >= 0; --k) {
			const unsigned c = node[t].branch[k];
			if (node[c].bodies) down = c;
		}
	unsigned next = kEnd;
	unsigned n = t;
	for (int up = 0; up <= kDepthCap && n != 0u; ++up) {
		const unsigned par = node[n].parent;
		int k = 0;
		while (k < 8 && node[par].branch[k] != n) ++k;
		for (++k; k < 8; ++k) {
			const unsigned c = node[par].branch[k];
			if (node[c].bodies) {
				next = c;
				break;
			}
		}
		if (next != kEnd) break;
		n = par;
	}
	link[t] = make_uint2(down, next);
}

}  // namespace

extern "C" int sfd_tree_bounds_ref(const float* pos, unsigned N, const double jitter[3], double* r1, float* bounds,
                                   hipStream_t s) {
	if (N == 0) return 0;
	hipLaunchKernelGGL(k_radius_f64, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, N, r1);
	hipLaunchKernelGGL(k_tree_bounds_ref, dim3(1), dim3(64), 0, s, r1, N, jitter[0], jitter[1], jitter[2], bounds);
	return sf_hip_check(hipGetLastError(), "sfd_tree_bounds_ref", "launch");
}

extern "C" int sfd_mean_radius(const float* pos, unsigned N, double* r1, double* avg, hipStream_t s) {
	if (N == 0) return 0;
	hipLaunchKernelGGL(k_radius_f64, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, N, r1);
	hipLaunchKernelGGL(k_mean_radius, dim3(1), dim3(64), 0, s, r1, N, avg);
	return sf_hip_check(hipGetLastError(), "sfd_mean_radius", "launch");
}

extern "C" int sfd_tree_build_ref(const float* pos, const float* mas, unsigned N, const float* bounds,
                                  unsigned max_nodes, int depth_max, sf_node* node, unsigned* meta, hipStream_t s) {
	const int tok = sf_prof_begin("tree_build_ref", s);
	hipLaunchKernelGGL(k_tree_build_ref, dim3(1), dim3(64), 0, s, pos, mas, N, bounds, max_nodes, depth_max, node,
	                   meta);
	if (sf_hip_check(hipGetLastError(), "sfd_tree_build_ref", "build launch")) return -1;
	const unsigned grid = (max_nodes + kBlock - 1) / kBlock;
	hipLaunchKernelGGL(k_tree_com_ref, dim3(grid ? grid : 1), dim3(kBlock), 0, s, node, meta);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_tree_build_ref", "com launch");
}

extern "C" int sfd_tree_walk_ref(const float* pos, const float* mas, unsigned N, const unsigned* order,
                                 const sf_node* node, unsigned node_cap, unsigned iter_max, float eps2, float theta, float negG, int acc_phys, int kahan_acc,
                                 int kahan_pot, float* acc, float* pot, hipStream_t s) {
	if (N == 0) return 0;
	const dim3 grid((N + kBlock - 1) / kBlock), block(kBlock);
	const int sel = (kahan_acc ? 4 : 0) | (kahan_pot ? 2 : 0) | (acc_phys ? 1 : 0);
	const int tok = sf_prof_begin("tree_walk_ref", s);
#define SF_WALK(A, B, C)                                                                                         \
	hipLaunchKernelGGL((k_tree_walk_ref<A, B, C>), grid, block, 0, s, pos, mas, N, order, node, node_cap, iter_max, eps2, theta, \
	                   negG, acc, pot)
	switch (sel) {
	case 0: SF_WALK(false, false, false); break;
	case 1: SF_WALK(false, false, true); break;
	case 2: SF_WALK(false, true, false); break;
	case 3: SF_WALK(false, true, true); break;
	case 4: SF_WALK(true, false, false); break;
	case 5: SF_WALK(true, false, true); break;
	case 6: SF_WALK(true, true, false); break;
	default: SF_WALK(true, true, true); break;
	}
#undef SF_WALK
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_tree_walk_ref", "walk launch");
}

extern "C" int sfd_tree_links(const sf_node* node, const unsigned* meta, unsigned max_nodes, void* link,
                              hipStream_t s) {
	const unsigned grid = (max_nodes + kBlock - 1) / kBlock;
	hipLaunchKernelGGL(k_tree_links, dim3(grid ? grid : 1), dim3(kBlock), 0, s, node, meta,
	                   static_cast<uint2*>(link));
	return sf_hip_check(hipGetLastError(), "sfd_tree_links", "launch");
}

extern "C" int sfd_tree_walk_thr(const float* pos, const float* mas, unsigned N, const unsigned* order,
                                 const sf_node* node, const void* link, float eps2, float theta, float negG,
                                 int acc_phys, int kahan_acc, int kahan_pot, float* acc, float* pot, hipStream_t s) {
	if (N == 0) return 0;
	const dim3 grid((N + kBlock - 1) / kBlock), block(kBlock);
	const uint2* ln = static_cast<const uint2*>(link);
	const int sel = (kahan_acc ? 4 : 0) | (kahan_pot ? 2 : 0) | (acc_phys ? 1 : 0);
	const int tok = sf_prof_begin("tree_walk_ref", s);
#define SF_WALK(A, B, C)                                                                                         \
	hipLaunchKernelGGL((k_tree_walk_thr<A, B, C>), grid, block, 0, s, pos, mas, N, order, node, ln, eps2, theta, negG, \
	                   acc, pot)
	switch (sel) {
	case 0: SF_WALK(false, false, false); break;
	case 1: SF_WALK(false, false, true); break;
	case 2: SF_WALK(false, true, false); break;
	case 3: SF_WALK(false, true, true); break;
	case 4: SF_WALK(true, false, false); break;
	case 5: SF_WALK(true, false, true); break;
	case 6: SF_WALK(true, true, false); break;
	default: SF_WALK(true, true, true); break;
	}
#undef SF_WALK
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_tree_walk_thr", "walk launch");
}
