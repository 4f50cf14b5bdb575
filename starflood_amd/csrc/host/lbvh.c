/*
 * lbvh.c -- host side of the SF_TREE_LBVH Barnes-Hut tree: device buffers
 * sized for N, the build sequence and the walk (kernels in device/lbvh.hip).
 */
#include <stdlib.h>
#include <string.h>

#include "../device/tree_dev.h"
#include "sf_host.h"

/* Walk group size cap: groups are the maximal subtrees with at most this many
   particles (one wavefront each). */
#define SF_LBVH_GROUP 64

struct sf_lbvh {
	unsigned int n_alloc;     /* N the buffers are sized for */
	unsigned int N;           /* particles in the last build */
	unsigned int n_sort;
	unsigned long long* key0;
	unsigned long long* key1;
	unsigned int* val0;
	unsigned int* val1;
	unsigned int* hist;
	double* bbox_part;
	double* box;
	void* body;               /* N float4, sorted (x, y, z, m) */
	void* node_lr;            /* N-1 int2 */
	void* node_range;         /* N-1 int2 */
	void* prefix;             /* N x 4 doubles: inclusive (m, mx, my, mz) sums */
	void* tile_sum;           /* scan tiles x 4 doubles */
	void* node_cm;            /* 2N-1 float4: com, mass */
	void* node_cell;          /* 2N-1 float4: Morton cell centre, side */
	void* rec_b;              /* 2N-1 int4 */
	unsigned long long* counters;   /* 6: the walk's four (zeroed per walk), the sticky guard count, the
	                                   walk's chunk counter (zeroed per walk) */
	unsigned long long* h_counters; /* pinned */
	int* gstack;                    /* walk node stacks, one per wavefront slot */
	unsigned int* group_start;      /* walk group starts in sorted order */
	/* quadrupoles (allocated on first use) */
	void* prefix2;                  /* N x 6 doubles: inclusive m x' x'^T sums */
	void* tile_sum2;
	double* quad_ref;               /* 3 doubles: reference point of the sums */
	void* node_q0;                  /* 2N-1 float4 */
	float* node_q1;                 /* 2N-1 floats */
	int quad_valid;                 /* the last build computed them */
};

static void free_buffers(sf_lbvh* t) {
	void* ptrs[] = {t->key0, t->key1, t->val0, t->val1, t->hist, t->bbox_part, t->box, t->body, t->node_lr,
	                t->node_range, t->prefix, t->tile_sum, t->node_cm, t->node_cell, t->rec_b, t->counters,
	                t->gstack, t->group_start, t->prefix2, t->tile_sum2, t->quad_ref, t->node_q0, t->node_q1};
	for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k)
		if (ptrs[k]) hipFree(ptrs[k]);
	if (t->h_counters) hipHostFree(t->h_counters);
	memset(t, 0, sizeof(*t));
}

void sf_lbvh_destroy(sf_lbvh* t) {
	if (!t) return;
	hipDeviceSynchronize();
	free_buffers(t);
	free(t);
}

static int alloc(void** p, size_t bytes) {
	if (hipMalloc(p, bytes ? bytes : 16u) == hipSuccess) return 0;
	*p = NULL;
	return -1;
}

static int ensure(sf_lbvh** st, unsigned int N) {
	if (!*st) {
		*st = (sf_lbvh*)calloc(1, sizeof(sf_lbvh));
		if (!*st) return -1;
	}
	sf_lbvh* t = *st;
	if (t->n_alloc >= N && t->key0) return 0;
	hipDeviceSynchronize();
	free_buffers(t);
	const unsigned int tile = sfd_radix_tile();
	const size_t n_sort = ((size_t)N + tile - 1u) / tile * tile;
	const size_t nn = 2u * (size_t)N;   /* >= 2N-1 */
	int rc = 0;
	rc |= alloc((void**)&t->key0, n_sort * 8u);
	rc |= alloc((void**)&t->key1, n_sort * 8u);
	rc |= alloc((void**)&t->val0, n_sort * 4u);
	rc |= alloc((void**)&t->val1, n_sort * 4u);
	rc |= alloc((void**)&t->hist, (size_t)sfd_radix_hist_entries((unsigned int)n_sort) * 4u);
	rc |= alloc((void**)&t->bbox_part, sfd_lbvh_bbox_blocks() * 6u * 8u);
	rc |= alloc((void**)&t->box, 4u * 8u);
	rc |= alloc(&t->body, (size_t)N * 16u);
	rc |= alloc(&t->node_lr, (size_t)N * 8u);
	rc |= alloc(&t->node_range, (size_t)N * 8u);
	rc |= alloc(&t->prefix, (size_t)N * 32u);
	rc |= alloc(&t->tile_sum, (size_t)sfd_lbvh_scan_tiles(N) * 32u);
	rc |= alloc(&t->node_cm, nn * 16u);
	rc |= alloc(&t->node_cell, nn * 16u);
	rc |= alloc(&t->rec_b, nn * 16u);
	rc |= alloc((void**)&t->counters, 6u * 8u);
	rc |= alloc((void**)&t->gstack, sfd_lbvh_walk_stack_bytes());
	rc |= alloc((void**)&t->group_start, (size_t)sfd_lbvh_group_flag_len(N) * 4u);
	if (!rc && hipHostMalloc((void**)&t->h_counters, 5u * 8u, hipHostMallocDefault) != hipSuccess) rc = -1;
	/* the sticky guard count starts at zero (the walk resets only the first four) */
	if (!rc && (hipMemset(t->counters, 0, 6u * 8u) != hipSuccess || hipDeviceSynchronize() != hipSuccess)) rc = -1;
	if (rc) {
		sf_error("tree_build", "device allocation for an LBVH of %u particles failed", N);
		free_buffers(t);
		return -1;
	}
	t->n_alloc = N;
	t->n_sort = (unsigned int)n_sort;
	return 0;
}

static void free_quad(sf_lbvh* t) {
	void** ptrs[] = {&t->prefix2, &t->tile_sum2, (void**)&t->quad_ref, &t->node_q0, (void**)&t->node_q1};
	for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) {
		if (*ptrs[k]) hipFree(*ptrs[k]);
		*ptrs[k] = NULL;
	}
}

/* The quadrupole buffers exist all together or not at all: a partial
   allocation is released, so the next build tries again from scratch. */
static int ensure_quad(sf_lbvh* t) {
	if (t->prefix2 && t->tile_sum2 && t->quad_ref && t->node_q0 && t->node_q1) return 0;
	free_quad(t);
	const size_t N = t->n_alloc, nn = 2u * N;
	int rc = 0;
	rc |= alloc(&t->prefix2, N * 48u);
	rc |= alloc(&t->tile_sum2, (size_t)sfd_lbvh_scan_tiles((unsigned int)N) * 48u);
	rc |= alloc((void**)&t->quad_ref, 3u * 8u);
	rc |= alloc(&t->node_q0, nn * 16u);
	rc |= alloc((void**)&t->node_q1, nn * 4u);
	if (rc) {
		sf_error("tree_build", "device allocation for LBVH quadrupoles of %zu particles failed", N);
		free_quad(t);
		return -1;
	}
	return 0;
}

int sf_lbvh_build(sf_lbvh** st, unsigned int N, const float* pos, const float* mas, int quadrupole, hipStream_t s) {
	if (N == 0) {
		if (*st) (*st)->N = 0;
		return 0;
	}
	if (ensure(st, N)) return -1;
	sf_lbvh* t = *st;
	t->quad_valid = 0;
	if (quadrupole && ensure_quad(t)) return -1;
	const unsigned int tile = sfd_radix_tile();
	t->n_sort = (N + tile - 1u) / tile * tile;
	t->N = N;
	const int tok = sf_prof_begin("lbvh_build", s);
	/* the prefix buffer holds the packed (x, y, z, m) records until the topology's scan */
	if (sfd_lbvh_keys(pos, mas, N, t->n_sort, t->bbox_part, t->box, t->key0, t->val0, t->prefix, s)) return -1;
	if (sfd_radix_sort(t->key0, t->val0, t->key1, t->val1, t->n_sort, t->hist, s)) return -1;
	if (sfd_lbvh_topology(t->key0, t->val0, N, t->box, t->body, t->node_lr, t->node_range, t->prefix, t->tile_sum,
	                      t->node_cm, t->node_cell, s))
		return -1;
	if (sfd_lbvh_groups(N, t->node_lr, t->node_range, SF_LBVH_GROUP, t->group_start, s)) return -1;
	if (sfd_lbvh_group_boxes(N, t->body, t->group_start, t->node_range, t->node_cell, s)) return -1;
	if (quadrupole) {
		if (sfd_lbvh_quadrupoles(N, t->body, t->node_range, t->prefix, t->box, t->prefix2, t->tile_sum2, t->quad_ref,
		                         t->node_q0, t->node_q1, s))
			return -1;
		t->quad_valid = 1;
	}
	sf_prof_end(tok, s);
	return 0;
}

/* prev_acc: acc holds each particle's previous acceleration, which the
   relative opening test (p->mac_alpha > 0) needs; without it the walk uses
   the theta test alone.  The records are made once per build. */
int sf_lbvh_walk(sf_lbvh* t, unsigned int N, const float* pos, const float* mas, const sf_params_t* p, int prev_acc,
                 float* acc, float* pot, hipStream_t s) {
	(void)pos;
	(void)mas;
	if (N == 0) return 0;
	if (!t || t->N != N) {
		sf_error("solve_gravity_part_tree", "the tree was built for %u particles, not %u", t ? t->N : 0u, N);
		return -1;
	}
	const int quad = p->quadrupole && t->quad_valid;
	if (p->quadrupole && !t->quad_valid) {
		sf_error("solve_gravity_part_tree", "quadrupole is set but the tree was built without quadrupoles");
		return -1;
	}
	if (sfd_lbvh_records(N, t->node_lr, t->node_range, t->node_cm, t->node_cell, (float)p->theta, (int)p->leaf_size,
	                     t->rec_b, s))
		return -1;
	/* G M l^p <= alpha |a| d^(p+2)  <=>  M l^p <= (alpha / G) |a| d^(p+2) */
	const float alpha_g = (prev_acc && p->mac_alpha > 0.0) ? (float)(p->mac_alpha / p->G) : 0.0f;
	return sfd_lbvh_walk(t->body, t->val0, N, t->group_start, t->node_cm, quad ? t->node_q0 : NULL,
	                     quad ? t->node_q1 : NULL, t->rec_b, alpha_g, (float)(p->epsilon * p->epsilon),
	                     p->G, p->acc_mode == SF_ACC_PHYSICAL, acc, pot, t->counters, t->gstack, s);
}

unsigned int sf_lbvh_size(const sf_lbvh* t) { return t ? t->N : 0u; }

/* Wavefronts of the walks since the last call that met a malformed tree or
   ran out of stack or node budget (their forces are incomplete); returns
   the count and clears it.  Synchronises s. */
int sf_lbvh_guard(sf_lbvh* t, hipStream_t s, unsigned long long* tripped) {
	*tripped = 0;
	if (!t || !t->counters) return 0;   /* nothing was walked */
	SF_HIP(hipMemcpyAsync(t->h_counters + 4, t->counters + 4, 8u, hipMemcpyDeviceToHost, s), "sf_lbvh_guard");
	SF_HIP(hipStreamSynchronize(s), "sf_lbvh_guard");
	*tripped = t->h_counters[4];
	if (*tripped) SF_HIP(hipMemsetAsync(t->counters + 4, 0, 8u, s), "sf_lbvh_guard");
	return 0;
}

int sf_lbvh_export_impl(sf_lbvh* t, hipStream_t s, unsigned long long* keys, unsigned int* order,
                        unsigned long long counts[4], float* node_cm, int* node_lr_range, float* node_cell) {
	if (!t || t->N == 0) {
		sf_error("sf_lbvh_export", "no LBVH has been built");
		return -1;
	}
	const size_t N = t->N;
	SF_HIP(hipStreamSynchronize(s), "sf_lbvh_export");
	if (keys) SF_HIP(hipMemcpy(keys, t->key0, N * 8u, hipMemcpyDeviceToHost), "sf_lbvh_export");
	if (order) SF_HIP(hipMemcpy(order, t->val0, N * 4u, hipMemcpyDeviceToHost), "sf_lbvh_export");
	if (counts) {
		SF_HIP(hipMemcpy(t->h_counters, t->counters, 4u * 8u, hipMemcpyDeviceToHost), "sf_lbvh_export");
		for (int k = 0; k < 4; ++k) counts[k] = t->h_counters[k];
	}
	if (node_cm) SF_HIP(hipMemcpy(node_cm, t->node_cm, (2u * N - 1u) * 16u, hipMemcpyDeviceToHost), "sf_lbvh_export");
	if (node_cell)
		SF_HIP(hipMemcpy(node_cell, t->node_cell, (2u * N - 1u) * 16u, hipMemcpyDeviceToHost), "sf_lbvh_export");
	if (node_lr_range && N > 1) {
		int* tmp = (int*)malloc((N - 1u) * 8u);
		if (!tmp) return -1;
		if (hipMemcpy(tmp, t->node_lr, (N - 1u) * 8u, hipMemcpyDeviceToHost) != hipSuccess) {
			free(tmp);
			return -1;
		}
		for (size_t i = 0; i + 1 < N; ++i) {
			node_lr_range[4 * i + 0] = tmp[2 * i + 0];
			node_lr_range[4 * i + 1] = tmp[2 * i + 1];
		}
		if (hipMemcpy(tmp, t->node_range, (N - 1u) * 8u, hipMemcpyDeviceToHost) != hipSuccess) {
			free(tmp);
			return -1;
		}
		for (size_t i = 0; i + 1 < N; ++i) {
			node_lr_range[4 * i + 2] = tmp[2 * i + 0];
			node_lr_range[4 * i + 3] = tmp[2 * i + 1];
		}
		free(tmp);
	}
	return 0;
}
