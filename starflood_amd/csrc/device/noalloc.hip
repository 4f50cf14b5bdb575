// noalloc.hip -- sf_force_dev_noalloc: the PP force on caller device arrays
// (sf_force_dev's kernels and summation order) with its scratch memory in a
// static device array instead of hipMalloc'd buffers, for callers that
// cannot allocate (code under stream capture, allocation-free serving
// loops).  No counterpart in the reference.
//
// scratch holds the packed bodies (n_pad float4, then the pack kernel's
// SF_BODY_STATE words) and, from the next 256-byte boundary, the slice
// partials (one double4 per particle and block slice).  It is sized for N up
// to SF_NOALLOC_MAX_N, where n_pad = N and the fast kernel writes 16 block
// slices: 256 KiB + 256 B of bodies and 8 MiB of partials.  Each device
// holds its own copy; its address comes from hipGetSymbolAddress, once per
// device, on the first call with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../sf_internal.h"

extern "C" {
size_t sf_pp_part_bytes(unsigned int n_pad);
int sf_pp_force_arrays(const sf_params_t* p, unsigned int N, unsigned int n_pad, const float* pos, const float* mas,
                       float* body, float* part, size_t part_bytes, float* acc, float* pot, hipStream_t s);
int sfd_fill0(float* y, unsigned int n, hipStream_t s);
}

namespace {

constexpr unsigned kMaxN = SF_NOALLOC_MAX_N;
constexpr unsigned kMaxBlockSlices = 16;   // of k_pp_fast at n_pad = kMaxN
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kBodyBytes = align_up(SF_BODY_BYTES(kMaxN));
constexpr size_t kPartBytes = (size_t)kMaxBlockSlices * kMaxN * 32u;

static __device__ __align__(256) unsigned char scratch[kBodyBytes + kPartBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char* g_scratch[kMaxDevices];   // device address of scratch, per device

// The scratch address on the current device, looked up on first use.
unsigned char* scratch_ptr(const char* fn) {
	int dev = -1;
	if (sf_hip_check(hipGetDevice(&dev), fn, "hipGetDevice")) return nullptr;
	if (dev < 0 || dev >= kMaxDevices) {
		sf_error(fn, "device %d is beyond the %d this entry point serves", dev, kMaxDevices);
		return nullptr;
	}
	std::lock_guard<std::mutex> hold(g_lock);
	if (!g_scratch[dev]) {
		void* p = nullptr;
		if (sf_hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), fn, "hipGetSymbolAddress(scratch)"))
			return nullptr;
		g_scratch[dev] = static_cast<unsigned char*>(p);
	}
	return g_scratch[dev];
}

}  // namespace

extern "C" int sf_force_dev_noalloc(unsigned int N, const float* x, const float* m, float* acc, float* pot,
                                    void* stream) {
	const char* fn = "sf_force_dev_noalloc";
	if (N == 0) return STARFLOOD_SUCCESS;
	if (!x || !m || !acc) {
		sf_error(fn, "null array");
		return STARFLOOD_FAILURE;
	}
	if (N > kMaxN) {
		sf_error(fn, "N = %u exceeds the static scratch's maximum of %u particles", N, kMaxN);
		return STARFLOOD_FAILURE;
	}
	// the layout sf_force_dev allocates: n_pad a multiple of 1024
	const unsigned n_pad = (N + 1023u) / 1024u * 1024u;
	const size_t body_bytes = align_up(SF_BODY_BYTES(n_pad));
	const size_t part_bytes = sf_pp_part_bytes(n_pad);
	if (body_bytes + part_bytes > sizeof(scratch)) {
		sf_error(fn, "N = %u needs %zu scratch bytes, more than the %zu available", N, body_bytes + part_bytes,
		         sizeof(scratch));
		return STARFLOOD_FAILURE;
	}
	unsigned char* base = scratch_ptr(fn);
	if (!base) return STARFLOOD_FAILURE;
	float* body = reinterpret_cast<float*>(base);
	float* part = reinterpret_cast<float*>(base + body_bytes);
	hipStream_t s = static_cast<hipStream_t>(stream);
	// The pack kernel's state words follow the n_pad bodies, so they move
	// with N and may hold an earlier call's bodies: clear them (a kernel, not
	// a memset node, so the call can be captured).
	if (sfd_fill0(body + 4u * (size_t)n_pad, SF_BODY_STATE, s)) return STARFLOOD_FAILURE;
	return sf_pp_force_arrays(sf_params(), N, n_pad, x, m, body, part, part_bytes, acc, pot, s) ? STARFLOOD_FAILURE
	                                                                                           : STARFLOOD_SUCCESS;
}
