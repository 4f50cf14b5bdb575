/*
 * context.c -- device state of each live simulation, keyed by sim_t.mem.
 *
 * A sim_t is a plain struct the caller copies by value (reference
 * src/simulation.c:78,184), so the library cannot hang its device state off
 * it; the host buffer address sim.mem identifies the simulation instead.
 * The reference keeps file-scope log handles and so supports one live
 * simulation per process; the registry allows a few, each with its own
 * stream, but calls on one simulation are not reentrant.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../sf_internal.h"
#include "sf_host.h"

#define SF_MAX_CTX 16

static sf_ctx* g_ctx[SF_MAX_CTX];
static hipStream_t g_default_stream = NULL;
static int g_default_stream_ready = 0;

hipStream_t sf_default_stream(void) {
	if (!g_default_stream_ready) {
		if (hipStreamCreateWithFlags(&g_default_stream, hipStreamNonBlocking) != hipSuccess) g_default_stream = NULL;
		g_default_stream_ready = 1;
	}
	return g_default_stream;
}

int sf_ctx_dof_conflict(unsigned int dof) {
	for (int k = 0; k < SF_MAX_CTX; ++k)
		if (g_ctx[k] && g_ctx[k]->dof != dof) return 1;
	return 0;
}

sf_ctx* sf_ctx_lookup(const void* host_mem, const char* func) {
	if (host_mem) {
		for (int k = 0; k < SF_MAX_CTX; ++k)
			if (g_ctx[k] && g_ctx[k]->host_mem == host_mem) return g_ctx[k];
	}
	sf_error(func, "sim.mem %p was not allocated by sim_init()", host_mem);
	return NULL;
}

static size_t round_up(size_t v, size_t m) { return (v + m - 1u) / m * m; }

sf_ctx* sf_ctx_create(void* host_mem, unsigned int N, unsigned int dof) {
	int slot = -1;
	for (int k = 0; k < SF_MAX_CTX; ++k)
		if (!g_ctx[k]) { slot = k; break; }
	if (slot < 0) {
		sf_error("sim_init", "more than %d live simulations", SF_MAX_CTX);
		return NULL;
	}
	sf_ctx* c = (sf_ctx*)calloc(1, sizeof(sf_ctx));
	if (!c) {
		sf_error("sim_init", "out of host memory");
		return NULL;
	}
	c->host_mem = host_mem;
	c->N = N;
	c->dof = dof;
	c->n_pad = (unsigned int)round_up(N ? N : 1u, 1024u);
	const size_t nm = (size_t)N * dof;
	if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) goto fail;
	if (hipMalloc((void**)&c->d_mem, (nm ? nm : 1u) * sizeof(float)) != hipSuccess) goto fail;
	if (hipMalloc((void**)&c->d_body, SF_BODY_BYTES(c->n_pad)) != hipSuccess) goto fail;
	if (hipMalloc((void**)&c->d_x64, ((size_t)N * 3u + 1u) * sizeof(double)) != hipSuccess) goto fail;
	if (hipMalloc((void**)&c->d_red, (size_t)sfd_reduce_scratch_doubles(N) * sizeof(double)) != hipSuccess) goto fail;
	if (hipHostMalloc((void**)&c->h_red, SF_RED_VALUES * sizeof(double), hipHostMallocDefault) != hipSuccess) goto fail;
	if (hipMemsetAsync(c->d_mem, 0, (nm ? nm : 1u) * sizeof(float), c->stream) != hipSuccess) goto fail;
	if (hipMemsetAsync(c->d_body, 0, SF_BODY_BYTES(c->n_pad), c->stream) != hipSuccess) goto fail;
	if (hipMemsetAsync(c->d_x64, 0, ((size_t)N * 3u + 1u) * sizeof(double), c->stream) != hipSuccess) goto fail;
	if (hipStreamSynchronize(c->stream) != hipSuccess) goto fail;
	g_ctx[slot] = c;
	return c;
fail:
	sf_error("sim_init", "device allocation for N = %u failed: %s", N, hipGetErrorString(hipGetLastError()));
	sf_ctx_destroy(c);
	return NULL;
}

void sf_ctx_destroy(sf_ctx* c) {
	if (!c) return;
	for (int k = 0; k < SF_MAX_CTX; ++k)
		if (g_ctx[k] == c) g_ctx[k] = NULL;
	if (c->stream) hipStreamSynchronize(c->stream);
	if (c->d_mem) hipFree(c->d_mem);
	if (c->d_body) hipFree(c->d_body);
	if (c->d_x64) hipFree(c->d_x64);
	if (c->d_part) hipFree(c->d_part);
	if (c->d_red) hipFree(c->d_red);
	if (c->h_red) hipHostFree(c->h_red);
	if (c->d_u) hipFree(c->d_u);
	if (c->d_dudt) hipFree(c->d_dudt);
	if (c->d_sph_tmp) hipFree(c->d_sph_tmp);
	if (c->d_sph_grid) hipFree(c->d_sph_grid);
	if (c->stream) hipStreamDestroy(c->stream);
	free(c);
}

int sf_ctx_reserve_part(sf_ctx* c, unsigned int slices) {
	if (c->d_part && c->part_slices >= slices) return 0;
	if (c->d_part) {
		hipStreamSynchronize(c->stream);
		hipFree(c->d_part);
		c->d_part = NULL;
	}
	const size_t bytes = (size_t)slices * 2u * 4u * c->n_pad * sizeof(float);
	SF_HIP(hipMalloc((void**)&c->d_part, bytes), "sf_ctx_reserve_part");
	c->part_slices = slices;
	return 0;
}

int sf_upload_blocks(sf_ctx* c, unsigned int e0, unsigned int count) {
	if (sf_params()->mem_mode != SF_MEM_MIRROR || c->N == 0 || count == 0) return 0;
	const size_t off = (size_t)c->N * e0, n = (size_t)c->N * count;
	SF_HIP(hipMemcpyAsync(c->d_mem + off, (float*)c->host_mem + off, n * sizeof(float), hipMemcpyHostToDevice,
	                      c->stream),
	       "sf_upload_blocks");
	return 0;
}

int sf_download_blocks(sf_ctx* c, unsigned int e0, unsigned int count) {
	if (sf_params()->mem_mode != SF_MEM_MIRROR || c->N == 0 || count == 0) return 0;
	const size_t off = (size_t)c->N * e0, n = (size_t)c->N * count;
	SF_HIP(hipMemcpyAsync((float*)c->host_mem + off, c->d_mem + off, n * sizeof(float), hipMemcpyDeviceToHost,
	                      c->stream),
	       "sf_download_blocks");
	return 0;
}

int sf_finish(sf_ctx* c, const char* func) {
	SF_HIP(hipStreamSynchronize(c->stream), func);
	SF_HIP(hipGetLastError(), func);
	return 0;
}

int sf_sync(const sim_t* sim_ptr, int direction) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_sync");
	if (!c) return STARFLOOD_FAILURE;
	const size_t n = (size_t)c->N * c->dof * sizeof(float);
	if (n) {
		if (direction == SF_SYNC_TO_DEVICE) {
			SF_HIP(hipMemcpyAsync(c->d_mem, c->host_mem, n, hipMemcpyHostToDevice, c->stream), "sf_sync");
		} else if (direction == SF_SYNC_TO_HOST) {
			SF_HIP(hipMemcpyAsync(c->host_mem, c->d_mem, n, hipMemcpyDeviceToHost, c->stream), "sf_sync");
		} else {
			sf_error("sf_sync", "unknown direction %d", direction);
			return STARFLOOD_FAILURE;
		}
	}
	int rc = sf_finish(c, "sf_sync");
	/* resident-mode results reach the host here: report a tree walk cut short */
	if (!rc && direction == SF_SYNC_TO_HOST && sf_params()->enable_tree && c->N &&
	    sf_tree_check((tree_t*)&sim_ptr->tree, c->stream, "sf_sync"))
		rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

void* sf_device_find(const sim_t* sim_ptr, enum sim_conf parameter_enum) {
	if (!sim_ptr) return NULL;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_device_find");
	if (!c) return NULL;
	unsigned int e = (unsigned int)parameter_enum;
	if (e >= c->dof) return NULL;
	if (e < 3u) e = SF_E_POS;
	else if (e < 6u) e = SF_E_VEL;
	else if (e < 9u) e = SF_E_ACC;
	return sf_blk(c->d_mem, c->N, e);
}

int sf_ctx_ensure_u(sf_ctx* c) {
	if (!c->d_u) {
		SF_HIP(hipMalloc((void**)&c->d_u, ((size_t)c->N + 1u) * sizeof(float)), "sf_device_ext");
		SF_HIP(hipMemsetAsync(c->d_u, 0, ((size_t)c->N + 1u) * sizeof(float), c->stream), "sf_device_ext");
	}
	if (!c->d_dudt) {
		SF_HIP(hipMalloc((void**)&c->d_dudt, ((size_t)c->N + 1u) * sizeof(float)), "sf_device_ext");
		SF_HIP(hipMemsetAsync(c->d_dudt, 0, ((size_t)c->N + 1u) * sizeof(float), c->stream), "sf_device_ext");
	}
	return 0;
}

void* sf_device_ext(const sim_t* sim_ptr, int which) {
	if (!sim_ptr) return NULL;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_device_ext");
	if (!c) return NULL;
	switch (which) {
	case SF_EXT_U:
		if (sf_ctx_ensure_u(c)) return NULL;
		c->have_u = 1;   /* the caller may write u through the pointer */
		return c->d_u;
	case SF_EXT_DUDT:
		if (sf_ctx_ensure_u(c)) return NULL;
		return c->d_dudt;
	case SF_EXT_X64: return c->d_x64;
	default: sf_error("sf_device_ext", "unknown array %d", which); return NULL;
	}
}

int sf_ext_copy(const sim_t* sim_ptr, int which, void* host, int direction) {
	void* d = sf_device_ext(sim_ptr, which);
	if (!d || !host) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "sf_ext_copy");
	const size_t bytes = which == SF_EXT_X64 ? (size_t)c->N * 3u * sizeof(double) : (size_t)c->N * sizeof(float);
	if (which == SF_EXT_U && direction == SF_SYNC_TO_DEVICE) c->have_u = 1;
	if (bytes) {
		if (direction == SF_SYNC_TO_DEVICE)
			SF_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream), "sf_ext_copy");
		else
			SF_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c->stream), "sf_ext_copy");
	}
	return sf_finish(c, "sf_ext_copy") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
