// vis.hip -- the reference's accumulation rasteriser (src/visualization.c:
// 500-736) on the device.
//
// One lane per (particle, sample): the particle's position through the
// model-view-projection matrix, its palette colour from the potential, a
// pcg4d-seeded Gaussian jitter (Box-Muller), and an integer RGBA splat into
// the accumulation buffer, all in the reference's FP32 operation order
// (this file is compiled with -ffp-contract=off).  SPATIAL_SAMPLES = 128
// puts the samples of one particle in two whole wavefronts, so a wavefront
// shares one colour: its lanes are grouped by pixel with ballots and each
// group adds (count x colour, count) with one set of atomics.  Integer sums
// do not depend on the order, so grouping changes no result.
//
// cosf, sinf and logf are the device library's.  They and glibc's are
// faithfully rounded but not identical, so a sample near a pixel edge, or a
// colour near an integer step of 100 x colour, can land differently from the
// reference's in rare cases.
#include <hip/hip_runtime.h>

#include "../host/rng.h"
#include "../sf_internal.h"

namespace {

constexpr unsigned kBlock = 256;
constexpr unsigned kSamples = 128;               // SPATIAL_SAMPLES, config.h:203
constexpr double kTau = 6.28318530717958647692528676655900577;

struct Mvp {
	float m[16];
};

__global__ __launch_bounds__(kBlock) void k_vis_clear(int* __restrict__ buf, size_t n) {
	const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (i < n) buf[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_vis_splat(const float* __restrict__ pos, const float* __restrict__ vel,
                                                      const float* __restrict__ pot, unsigned N, Mvp mvp,
                                                      unsigned step, float dt, unsigned sizex, unsigned sizey,
                                                      int* __restrict__ buf) {
	const size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x;
	const unsigned idx = (unsigned)(g / kSamples), samp = (unsigned)(g % kSamples);
	bool in_frame = false;
	unsigned pixel = 0u;
	int col[3] = {0, 0, 0};
	if (idx < N) {
		// src/visualization.c:506-548: homogeneous position through the MVP
		const float x[4] = {pos[3u * idx], pos[3u * idx + 1u], pos[3u * idx + 2u], 1.0f};
		const float v[3] = {vel[3u * idx], vel[3u * idx + 1u], vel[3u * idx + 2u]};
		float p[4];
		for (int i = 0; i < 4; ++i)
			p[i] = (((mvp.m[4 * i + 0] * x[0]) + (mvp.m[4 * i + 1] * x[1])) + (mvp.m[4 * i + 2] * x[2])) +
			       (mvp.m[4 * i + 3] * x[3]);
		// src/visualization.c:550-556: palette
		const float phase = 1.000e-1f * pot[idx];
		const float tau = (float)kTau;
		const float shift[3] = {(float)(kTau * (0.0 / 3.0)), (float)(kTau * (1.0 / 3.0)), (float)(kTau * (2.0 / 3.0))};
		for (int k = 0; k < 3; ++k) col[k] = (int)(100.0f * (0.5f * cosf(tau * phase - shift[k]) + 0.5f));
		// src/visualization.c:571-625: jitter and (zero) motion-blur offset
		uint32_t s[4] = {0xB79ABC95u + samp, 0xE0DA3F84u + idx, 0xAB75F07Bu + step, 0xCF52CA01u};
		sf_pcg4d(s);
		float r[4];
		for (int k = 0; k < 4; ++k) r[k] = (float)SF_INV_U32 * (float)s[k];
		const float n0 = sqrtf(-2.0f * logf(r[2])) * cosf(tau * r[3]);
		const float n1 = sqrtf(-2.0f * logf(r[2])) * sinf(tau * r[3]);
		const float so[2] = {0.750f * n0, 0.750f * n1};
		float time_offset = 0.0f;
		time_offset *= dt;
		float pi[3];
		for (int k = 0; k < 3; ++k) pi[k] = p[k] + time_offset * v[k];
		// src/visualization.c:689-702: pixel, truncated toward zero
		const float fx = ((so[0] + -0.5f) + ((0.5f * (float)sizey) * pi[0])) + (0.5f * (float)sizex);
		const float fy = ((so[1] + -0.5f) + ((0.5f * (float)sizey) * pi[1])) + (0.5f * (float)sizey);
		// a NaN coordinate converts to INT_MIN on the reference's x86 host:
		// outside the frame
		if (fx == fx && fy == fy) {
			const int cx = (int)fx, cy = (int)fy;
			if (0 <= cx && cx < (int)sizex && 0 <= cy && cy < (int)sizey) {
				in_frame = true;
				pixel = sizex * (unsigned)cy + (unsigned)cx;
			}
		}
	}
	// one atomic set per distinct pixel of the wavefront
	unsigned long long active = __ballot(in_frame);
	const unsigned lane = threadIdx.x & 63u;
	while (active) {
		const int leader = __ffsll((long long)active) - 1;
		const unsigned p = (unsigned)__shfl((int)pixel, leader);
		const unsigned long long same = __ballot(in_frame && pixel == p) & active;
		if ((int)lane == leader) {
			const int cnt = __popcll(same);
			int* px = buf + 4u * (size_t)p;
			atomicAdd(&px[0], cnt * col[0]);
			atomicAdd(&px[1], cnt * col[1]);
			atomicAdd(&px[2], cnt * col[2]);
			atomicAdd(&px[3], cnt);
		}
		active &= ~same;
	}
}

// src/visualization.c:714-736: scale, tanh tone map, alpha 1
__global__ __launch_bounds__(kBlock) void k_vis_finish(const int* __restrict__ buf, size_t npix, double scale,
                                                       float* __restrict__ render) {
	const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (i >= npix) return;
	for (int j = 0; j < 3; ++j) render[4u * i + j] = (float)tanh(scale * (double)buf[4u * i + j]);
	render[4u * i + 3u] = 1.0f;
}

}  // namespace

extern "C" int sfd_vis_draw(const float* pos, const float* vel, const float* pot, unsigned N, const float mvp[16],
                            unsigned step, float dt, unsigned sizex, unsigned sizey, double scale, int* buf,
                            float* render, hipStream_t s) {
	const size_t npix = (size_t)sizex * sizey;
	if (npix == 0) return 0;
	Mvp m;
	for (int k = 0; k < 16; ++k) m.m[k] = mvp[k];
	const int tok = sf_prof_begin("vis_draw", s);
	hipLaunchKernelGGL(k_vis_clear, dim3((unsigned)((4u * npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, buf,
	                   4u * npix);
	if (N) {
		const size_t lanes = (size_t)N * kSamples;
		hipLaunchKernelGGL(k_vis_splat, dim3((unsigned)((lanes + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, pos, vel,
		                   pot, N, m, step, dt, sizex, sizey, buf);
	}
	hipLaunchKernelGGL(k_vis_finish, dim3((unsigned)((npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, buf, npix,
	                   scale, render);
	sf_prof_end(tok, s);
	return sf_hip_check(hipGetLastError(), "vis_draw", "launch");
}
