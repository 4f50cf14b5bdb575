// pp.hip -- particle-particle gravity on gfx950.
//
// Two kernels compute the same quantity as the reference's
// solve_gravity_part_part (src/gravity.c:666-819):
//   sum_i = sum_j m_j r_ij / (r_ij^2 + eps^2)^(3/2),  r_ij = r_i - r_j
//   acc_i = -G m_i sum_i   (SF_ACC_REFERENCE)   or  -G sum_i (SF_ACC_PHYSICAL)
//   pot_i = -G sum_j m_j / sqrt(r_ij^2 + eps^2)
// with the self pair included (its force term is zero, its potential term is
// -G m_i / eps), as the reference does.
//
// k_pp_exact keeps the reference's FP32 operation order (no contraction: the
// file is compiled with -ffp-contract=off; IEEE divide and square root;
// per-pair Kahan; j in index order), so its output is bit-identical to the
// CPU path.  One lane owns one i; the j stream is wavefront-uniform, so the
// compiler fetches it with scalar loads into SGPRs.
//
// k_pp_fast is the throughput kernel: 13 VALU operations plus one v_rsq_f32
// per interaction, one i per lane, j fetched with scalar loads, partial sums
// per tile of SF_TJ interactions folded into a Kahan-compensated running
// sum, and the j range split into slices so the grid fills all 256 CUs.
// When every particle has the same mass m (k_pack_bodies checks the bits of
// each mass on the device), the whole tiles of real particles drop the
// per-pair mass product: they sum 1/s^3 (r_j - r_i) and 1/s with 12 VALU
// operations plus the rsqrt, and the tile partial is scaled by m before the
// fold.  A block's four wavefronts hold the same 64 particles and split its
// slice into four sub-slices; their sums meet in LDS in FP64, in wavefront
// order, so a block writes one double4 per particle.  Slice s is processed
// by blocks with blockIdx % slices == s, which the dispatcher places on XCD
// (s mod 8), so each XCD's L2 holds only its own part of the body array.
// A second kernel adds the slice partials per i in fixed order in FP64,
// which keeps the result deterministic.
#include <hip/hip_runtime.h>

#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned SF_TJ = 128;   // interactions per plain-FP32 partial sum
constexpr int kPrefetch = 4;      // bodies per scalar-load group (one s_load_dwordx16)
static_assert(SF_TJ % kPrefetch == 0, "tiles hold whole prefetch groups");

// One step of Kahan summation (reference src/gravity.c:771-784):
// y = x - c; t = s + y; c = (t - s) - y; s = t.
__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
	const float y = x - c;
	const float t = s + y;
	const float z = t - s;
	c = z - y;
	s = t;
}

template <bool KAHAN_ACC, bool KAHAN_POT, bool PHYS>
__global__ __launch_bounds__(kBlock) void k_pp_exact(const float* __restrict__ pos, const float* __restrict__ mas,
                                                     unsigned N, float eps2, float negG,
                                                     float* __restrict__ acc, float* __restrict__ pot) {
	const unsigned i = blockIdx.x * kBlock + threadIdx.x;
	const bool live = i < N;
	const unsigned ii = live ? i : N - 1u;
	const float xi = pos[3u * ii + 0u];
	const float yi = pos[3u * ii + 1u];
	const float zi = pos[3u * ii + 2u];
	const float mi = mas[ii];

	float sx = 0.0f, sy = 0.0f, sz = 0.0f, sp = 0.0f;
	float cx = 0.0f, cy = 0.0f, cz = 0.0f, cp = 0.0f;

	for (unsigned j = 0; j < N; ++j) {
		const float mj = mas[j];
		const float rx = xi - pos[3u * j + 0u];
		const float ry = yi - pos[3u * j + 1u];
		const float rz = zi - pos[3u * j + 2u];
		const float r2 = (rx * rx + ry * ry) + rz * rz;
		const float inv_r2 = 1.0f / (r2 + eps2);
		const float inv_r1 = 1.0f / sqrtf(r2 + eps2);
		const float pj = mj * inv_r1;
		const float fx = pj * inv_r2 * rx;
		const float fy = pj * inv_r2 * ry;
		const float fz = pj * inv_r2 * rz;
		if (KAHAN_ACC) {
			kahan_add(sx, cx, fx);
			kahan_add(sy, cy, fy);
			kahan_add(sz, cz, fz);
		} else {
			sx += fx;
			sy += fy;
			sz += fz;
		}
		if (KAHAN_POT) {
			kahan_add(sp, cp, pj);
		} else {
			sp += pj;
		}
	}
	if (live) {
		// reference src/gravity.c:808-811: ((-G) * m_i) * acc_sum, (-G) * pot_sum
		const float f = PHYS ? negG : negG * mi;
		acc[3u * i + 0u] = f * sx;
		acc[3u * i + 1u] = f * sy;
		acc[3u * i + 2u] = f * sz;
		pot[i] = negG * sp;
	}
}

// One tile of SF_TJ interactions, j from jt.  nb holds the tile's first
// kPrefetch bodies on entry and the next tile's first ones on exit (the
// group after the range's last is its own first again: in bounds; the
// compiler then issues the four loads as separate dwordx4 into fresh
// registers, early in the group -- measured faster than one dwordx16 into
// the registers being consumed).  UNIFORM: every mass of the tile equals m,
// and the sums leave it out (sum_j (r_j - r_i)/s^3 and sum_j 1/s).
template <bool WITH_POT, bool UNIFORM>
__device__ __forceinline__ void pp_tile(const float4* __restrict__ body, unsigned jt, unsigned j1,
                                        float4 (&nb)[kPrefetch], float xi, float yi, float zi, float eps2, float& tx,
                                        float& ty, float& tz, float& tp) {
	for (unsigned j = jt; j < jt + SF_TJ; j += kPrefetch) {
		float4 cb[kPrefetch];
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) cb[u] = nb[u];
		const unsigned jn = (j + kPrefetch < j1) ? j + kPrefetch : j;
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) nb[u] = body[jn + u];
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) {
			const float4 bj = cb[u];
			const float dx = bj.x - xi;
			const float dy = bj.y - yi;
			const float dz = bj.z - zi;
			float r2 = __builtin_fmaf(dx, dx, eps2);
			r2 = __builtin_fmaf(dy, dy, r2);
			r2 = __builtin_fmaf(dz, dz, r2);
			const float inv = __builtin_amdgcn_rsqf(r2);
			const float mi = UNIFORM ? inv : bj.w * inv;
			const float inv2 = inv * inv;
			const float w = mi * inv2;
			tx = __builtin_fmaf(w, dx, tx);
			ty = __builtin_fmaf(w, dy, ty);
			tz = __builtin_fmaf(w, dz, tz);
			if (WITH_POT) tp += mi;
		}
	}
}

// Block b owns the 64 particles i of i-block b / slices (one per lane, the
// same in each of its kWavesPP wavefronts) and j slice b % slices, which its
// wavefronts split into kWavesPP consecutive sub-slices of sub_len bodies.
// Each wavefront accumulates sum_j m_j (r_j - r_i) / s^3 and sum_j m_j / s
// over its sub-slice: plain FP32 sums over tiles of SF_TJ interactions,
// folded into Kahan-compensated running sums.  The wavefronts' compensated
// sums (s - c) are then added in FP64 in wavefront order through LDS, and
// the block writes one double4 per particle: part[slice][i].  The word
// after the n_pad bodies holds n_uni: the bodies [0, n_uni) are whole tiles
// of particles of one mass, body[0].w (0 when the masses differ).
constexpr unsigned kWavesPP = kBlock / 64;
template <bool WITH_POT>
__global__ __launch_bounds__(kBlock) void k_pp_fast(const float4* __restrict__ body, unsigned n_pad,
                                                    unsigned slices, unsigned sub_len, float eps2,
                                                    double4* __restrict__ part) {
	__shared__ double red[kWavesPP][4][64];
	const unsigned b = blockIdx.x;
	const unsigned s = b % slices;
	// the wavefront index is uniform per wavefront: readfirstlane tells the
	// compiler, so the j stream stays in scalar loads and SGPRs
	const unsigned lane = threadIdx.x & 63u, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const unsigned i = (b / slices) * 64u + lane;
	const float4 bi = body[i];
	const float xi = bi.x, yi = bi.y, zi = bi.z;
	float Sx = 0.0f, Sy = 0.0f, Sz = 0.0f, Sp = 0.0f;
	float Cx = 0.0f, Cy = 0.0f, Cz = 0.0f, Cp = 0.0f;

	// The j stream keeps one group of kPrefetch bodies in flight: the scalar
	// loads of the next group are issued before the current group's
	// arithmetic, so their latency overlaps it.  The equal-mass tiles come
	// first in a loop of their own: one loop with a branch per tile made the
	// compiler carry the j stream in VGPRs (measured 8 % slower).
	const unsigned j0 = min((s * kWavesPP + wv) * sub_len, n_pad);
	const unsigned j1 = min(j0 + sub_len, n_pad);
	if (j0 < j1) {
		const unsigned n_uni = *reinterpret_cast<const unsigned*>(body + n_pad);
		const unsigned ju = min(j1, max(j0, n_uni));
		const float mu = body[0].w;
		float4 nb[kPrefetch];
#pragma unroll
		for (int u = 0; u < kPrefetch; ++u) nb[u] = body[j0 + u];
		unsigned jt = j0;
		for (; jt < ju; jt += SF_TJ) {
			float tx = 0.0f, ty = 0.0f, tz = 0.0f, tp = 0.0f;
			pp_tile<WITH_POT, true>(body, jt, j1, nb, xi, yi, zi, eps2, tx, ty, tz, tp);
			kahan_add(Sx, Cx, mu * tx);
			kahan_add(Sy, Cy, mu * ty);
			kahan_add(Sz, Cz, mu * tz);
			if (WITH_POT) kahan_add(Sp, Cp, mu * tp);
		}
		for (; jt < j1; jt += SF_TJ) {
			float tx = 0.0f, ty = 0.0f, tz = 0.0f, tp = 0.0f;
			pp_tile<WITH_POT, false>(body, jt, j1, nb, xi, yi, zi, eps2, tx, ty, tz, tp);
			kahan_add(Sx, Cx, tx);
			kahan_add(Sy, Cy, ty);
			kahan_add(Sz, Cz, tz);
			if (WITH_POT) kahan_add(Sp, Cp, tp);
		}
	}
	// Kahan's compensation c holds the negated lost low part: each sum is s - c
	red[wv][0][lane] = (double)Sx - (double)Cx;
	red[wv][1][lane] = (double)Sy - (double)Cy;
	red[wv][2][lane] = (double)Sz - (double)Cz;
	red[wv][3][lane] = (double)Sp - (double)Cp;
	__syncthreads();
	if (wv != 0) return;
	double4 t = make_double4(red[0][0][lane], red[0][1][lane], red[0][2][lane], red[0][3][lane]);
#pragma unroll
	for (unsigned q = 1; q < kWavesPP; ++q) {
		t.x += red[q][0][lane];
		t.y += red[q][1][lane];
		t.z += red[q][2][lane];
		t.w += red[q][3][lane];
	}
	part[(size_t)s * n_pad + i] = t;
}

// Fixed-order FP64 sum of the slice partials.  A block takes kCombI
// particles; its wavefronts take the slices s = g, g + kCombG, ... (g =
// wavefront), each lane one particle (coalesced loads), and wavefront 0
// adds the kCombG sums in order.  One lane per particle over all slices
// left most CUs idle at small N (64 blocks at N = 16384: 20 us).
constexpr unsigned kCombI = 64;
constexpr unsigned kCombG = kBlock / kCombI;
__global__ __launch_bounds__(kBlock) void k_pp_combine(const double4* __restrict__ part, unsigned n_pad,
                                                       unsigned slices, const float4* __restrict__ body,
                                                       unsigned N, double G, int acc_phys,
                                                       float* __restrict__ acc, float* __restrict__ pot) {
	__shared__ double red[kCombG][4][kCombI];
	const unsigned il = threadIdx.x % kCombI, g = threadIdx.x / kCombI;
	const unsigned i = blockIdx.x * kCombI + il;   // < n_pad: the grid covers N in kCombI steps
	double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
	for (unsigned s = g; s < slices; s += kCombG) {
		const double4 a = part[(size_t)s * n_pad + i];
		sx += a.x;
		sy += a.y;
		sz += a.z;
		sp += a.w;
	}
	red[g][0][il] = sx;
	red[g][1][il] = sy;
	red[g][2][il] = sz;
	red[g][3][il] = sp;
	__syncthreads();
	if (g != 0 || i >= N) return;
	for (unsigned q = 1; q < kCombG; ++q) {
		sx += red[q][0][il];
		sy += red[q][1][il];
		sz += red[q][2][il];
		sp += red[q][3][il];
	}
	// sum_j m_j (r_j - r_i)/s^3 = -sum_j m_j r_ij/s^3, so acc = G (m_i) * sum
	const double f = acc_phys ? G : G * (double)body[i].w;
	acc[3u * i + 0u] = (float)(f * sx);
	acc[3u * i + 1u] = (float)(f * sy);
	acc[3u * i + 2u] = (float)(f * sz);
	if (pot) pot[i] = (float)(-G * sp);
}

// Packs (x, y, z, m) per body, padding with zeros to n_pad, and finds
// whether every mass has the bits of mas[0].  st[0] receives n_uni, the
// whole SF_TJ tiles of [0, N) when they do, else 0 (padding bodies, mass 0,
// are never in a uniform tile).  st[1] collects the blocks' differing-mass
// flags and st[2] counts finished blocks; the last block to finish writes
// st[0] and returns both to zero (atomicInc wraps at gridDim.x - 1), so
// they start at zero on the next call without a clear.  The grid is capped
// at kPackBlocks and strides over the bodies: the atomics on st[2] serialise
// (one per block of 1024 measured 26 us at N = 262144).
constexpr unsigned kPackBlocks = 128;
__global__ __launch_bounds__(kBlock) void k_pack_bodies(const float* __restrict__ pos, const float* __restrict__ mas,
                                                        unsigned N, unsigned n_pad, float4* __restrict__ body,
                                                        unsigned* __restrict__ st) {
	const unsigned m0 = N ? __float_as_uint(mas[0]) : 0u;
	int d = 0;
	for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < n_pad; i += gridDim.x * kBlock) {
		if (i < N) {
			const float m = mas[i];
			body[i] = make_float4(pos[3u * i], pos[3u * i + 1u], pos[3u * i + 2u], m);
			d |= __float_as_uint(m) != m0;
		} else {
			body[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		}
	}
	d = __syncthreads_or(d);
	if (threadIdx.x == 0) {
		if (d) atomicOr(&st[1], 1u);
		__threadfence();
		if (atomicInc(&st[2], gridDim.x - 1u) == gridDim.x - 1u) {
			const unsigned any = atomicExch(&st[1], 0u);
			st[0] = any ? 0u : N / SF_TJ * SF_TJ;
		}
	}
}

// FP64 direct sum: one block per listed particle, lanes stride over j, then
// a fixed-shape tree reduction in LDS.
__global__ __launch_bounds__(kBlock) void k_pp_f64(const float* __restrict__ pos, const float* __restrict__ mas,
                                                   unsigned N, double eps2, double G, int acc_phys,
                                                   const unsigned* __restrict__ idx,
                                                   double* __restrict__ acc, double* __restrict__ pot) {
	__shared__ double red[4][kBlock];
	const unsigned i = idx[blockIdx.x];
	const double xi = pos[3u * i], yi = pos[3u * i + 1u], zi = pos[3u * i + 2u];
	double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
	for (unsigned j = threadIdx.x; j < N; j += kBlock) {
		const double rx = xi - (double)pos[3u * j];
		const double ry = yi - (double)pos[3u * j + 1u];
		const double rz = zi - (double)pos[3u * j + 2u];
		const double s2 = rx * rx + ry * ry + rz * rz + eps2;
		const double inv1 = 1.0 / sqrt(s2);
		const double pj = (double)mas[j] * inv1;
		const double w = pj * inv1 * inv1;
		sx += w * rx;
		sy += w * ry;
		sz += w * rz;
		sp += pj;
	}
	red[0][threadIdx.x] = sx;
	red[1][threadIdx.x] = sy;
	red[2][threadIdx.x] = sz;
	red[3][threadIdx.x] = sp;
	__syncthreads();
	for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
		if (threadIdx.x < h) {
#pragma unroll
			for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const double f = acc_phys ? -G : -G * (double)mas[i];
		acc[3u * blockIdx.x + 0u] = f * red[0][0];
		acc[3u * blockIdx.x + 1u] = f * red[1][0];
		acc[3u * blockIdx.x + 2u] = f * red[2][0];
		pot[blockIdx.x] = -G * red[3][0];
	}
}


}  // namespace

extern "C" int sfd_pack_bodies(const float* pos, const float* mas, unsigned N, unsigned n_pad, float* body,
                               hipStream_t s) {
	const unsigned grid = min((n_pad + kBlock - 1) / kBlock, kPackBlocks);
	unsigned* st = reinterpret_cast<unsigned*>(body + 4u * (size_t)n_pad);
	hipLaunchKernelGGL(k_pack_bodies, dim3(grid), dim3(kBlock), 0, s, pos, mas, N, n_pad,
	                   reinterpret_cast<float4*>(body), st);
	return sf_hip_check(hipGetLastError(), "sfd_pack_bodies", "k_pack_bodies launch");
}

extern "C" int sfd_pp_exact(const float* pos, const float* mas, unsigned N, float eps2, float negG, int acc_phys,
                            int kahan_acc, int kahan_pot, float* acc, float* pot, hipStream_t s) {
	if (N == 0) return 0;
	const dim3 grid((N + kBlock - 1) / kBlock), block(kBlock);
	const int sel = (kahan_acc ? 4 : 0) | (kahan_pot ? 2 : 0) | (acc_phys ? 1 : 0);
	const int tok = sf_prof_begin("pp_exact", s);
	switch (sel) {
	case 0: hipLaunchKernelGGL((k_pp_exact<false, false, false>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 1: hipLaunchKernelGGL((k_pp_exact<false, false, true>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 2: hipLaunchKernelGGL((k_pp_exact<false, true, false>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 3: hipLaunchKernelGGL((k_pp_exact<false, true, true>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 4: hipLaunchKernelGGL((k_pp_exact<true, false, false>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 5: hipLaunchKernelGGL((k_pp_exact<true, false, true>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	case 6: hipLaunchKernelGGL((k_pp_exact<true, true, false>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	default: hipLaunchKernelGGL((k_pp_exact<true, true, true>), grid, block, 0, s, pos, mas, N, eps2, negG, acc, pot); break;
	}
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_pp_exact", "k_pp_exact launch");
}

// The slice geometry of k_pp_fast for n_pad bodies and `slices` j ranges
// (whole SF_TJ tiles, every one non-empty): sub-slice length, and block
// slices (kWavesPP sub-slices each), one double4 plane each.
struct PPGeometry {
	unsigned sub_len, bslices;
};
static PPGeometry pp_geometry(unsigned n_pad, unsigned slices) {
	const unsigned tiles = n_pad / SF_TJ;
	// a multiple of the wavefronts per block: a wavefront without a j range
	// would wait at the block's barrier and hold its slot (at N = 4M, two
	// ranges left half of every block idle: the force took 6.7 s, not 4.5 s)
	slices = (max(slices, 1u) + kWavesPP - 1u) / kWavesPP * kWavesPP;
	if (slices > tiles) slices = tiles;
	if (slices < 1u) slices = 1u;
	PPGeometry g;
	g.sub_len = (tiles + slices - 1) / slices * SF_TJ;
	const unsigned subs = (n_pad + g.sub_len - 1) / g.sub_len;
	g.bslices = (subs + kWavesPP - 1) / kWavesPP;
	return g;
}

extern "C" size_t sfd_pp_part_bytes(unsigned n_pad, unsigned slices) {
	return n_pad ? (size_t)pp_geometry(n_pad, slices).bslices * n_pad * sizeof(double4) : 0u;
}

extern "C" int sfd_pp_fast(const float* body, unsigned N, unsigned n_pad, float eps2, double G, int acc_phys,
                           float* part, unsigned slices, float* acc, float* pot, hipStream_t s) {
	if (N == 0) return 0;
	if (n_pad % 64u != 0 || n_pad % SF_TJ != 0) {
		sf_error("sfd_pp_fast", "n_pad %u is not a multiple of %u", n_pad, SF_TJ);
		return -1;
	}
	// part holds sfd_pp_part_bytes(n_pad, slices): one double4 plane per block slice
	const PPGeometry g = pp_geometry(n_pad, slices);
	const unsigned grid = (n_pad / 64u) * g.bslices;
	double4* p4 = reinterpret_cast<double4*>(part);
	const float4* b4 = reinterpret_cast<const float4*>(body);
	const int tok = sf_prof_begin("pp_fast", s);
	if (pot)
		hipLaunchKernelGGL(k_pp_fast<true>, dim3(grid), dim3(kBlock), 0, s, b4, n_pad, g.bslices, g.sub_len, eps2, p4);
	else
		hipLaunchKernelGGL(k_pp_fast<false>, dim3(grid), dim3(kBlock), 0, s, b4, n_pad, g.bslices, g.sub_len, eps2,
		                   p4);
	sf_prof_end(tok, s);
	if (sf_hip_check(hipGetLastError(), "sfd_pp_fast", "k_pp_fast launch")) return -1;
	const int tok2 = sf_prof_begin("pp_combine", s);
	hipLaunchKernelGGL(k_pp_combine, dim3((N + kCombI - 1) / kCombI), dim3(kBlock), 0, s, p4, n_pad, g.bslices, b4,
	                   N, G, acc_phys, acc, pot);
	sf_prof_end(tok2, s);
	return sf_hip_check(hipGetLastError(), "sfd_pp_fast", "k_pp_combine launch");
}

extern "C" int sfd_pp_f64(const float* pos, const float* mas, unsigned N, double eps2, double G, int acc_phys,
                          const unsigned* idx, unsigned n_idx, double* acc, double* pot, hipStream_t s) {
	if (n_idx == 0) return 0;
	const int tok = sf_prof_begin("pp_f64", s);
	hipLaunchKernelGGL(k_pp_f64, dim3(n_idx), dim3(kBlock), 0, s, pos, mas, N, eps2, G, acc_phys, idx, acc, pot);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "sfd_pp_f64", "k_pp_f64 launch");
}
