/*
 * tree.c -- tree_init / tree_free / tree_build and solve_gravity_part_tree
 * with the reference's signatures (src/tree.h:38-44, src/gravity.h:17).
 *
 * tree_t.node keeps the reference's meaning: a host array of node_t.  In
 * SF_TREE_REFERENCE mode it receives the device build after each build (in
 * SF_MEM_MIRROR mode); in SF_TREE_LBVH mode the device tree has a different
 * shape and tree_t.node is not written (num_nodes reports its node count).
 * The device side of a tree is found through the address of tree_t.node.
 */
#define _POSIX_C_SOURCE 200112L

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../device/tree_dev.h"
#include "rng.h"
#include "sf_host.h"

#define SF_MAX_TREES 16

struct sf_tree_dev {
	void* key;               /* tree_t.node (host) */
	unsigned int max_nodes;
	sf_node* d_node;         /* reference build: max_nodes nodes */
	float* d_bounds;         /* 6 floats */
	unsigned int* d_meta;    /* num_nodes, status */
	unsigned int* h_meta;    /* pinned mirror */
	float* d_stage;          /* staging for host pos (3N) and mas (N) */
	unsigned int stage_n;
	hipStream_t stream;      /* for tree_build called directly */
	sf_tree_par_ctx par;     /* level-synchronous reference build: scratch, */
	size_t par_bytes;        /* pinned state, side stream and events */
	double* d_r1;            /* radii for the reference bounds */
	unsigned int r1_n;
	unsigned int* d_order;   /* walk order of the last parallel build (N entries) */
	unsigned int order_n;    /* its N, 0 when the last build was serial */
	void* d_link;            /* threaded links of the last complete parallel build (max_nodes x 2 u32) */
	sf_lbvh* lbvh;           /* SF_TREE_LBVH state (lbvh.c) */
};

static sf_tree_dev* g_tree[SF_MAX_TREES];

static sf_tree_dev* tree_dev_lookup(const tree_t* t, const char* func) {
	if (t && t->node) {
		for (int k = 0; k < SF_MAX_TREES; ++k)
			if (g_tree[k] && g_tree[k]->key == (void*)t->node) return g_tree[k];
	}
	sf_error(func, "tree_t.node %p was not allocated by tree_init()", t ? (void*)t->node : NULL);
	return NULL;
}

static void tree_dev_destroy(sf_tree_dev* d) {
	if (!d) return;
	for (int k = 0; k < SF_MAX_TREES; ++k)
		if (g_tree[k] == d) g_tree[k] = NULL;
	if (d->stream) hipStreamSynchronize(d->stream);
	if (d->lbvh) sf_lbvh_destroy(d->lbvh);
	if (d->d_node) hipFree(d->d_node);
	if (d->d_bounds) hipFree(d->d_bounds);
	if (d->d_meta) hipFree(d->d_meta);
	if (d->h_meta) hipHostFree(d->h_meta);
	if (d->d_stage) hipFree(d->d_stage);
	if (d->par.side) hipStreamSynchronize(d->par.side);
	if (d->par.scratch) hipFree(d->par.scratch);
	if (d->par.h_state) hipHostFree(d->par.h_state);
	for (int k = 0; k < 2; ++k)
		if (d->par.ev[k]) hipEventDestroy(d->par.ev[k]);
	if (d->par.side) hipStreamDestroy(d->par.side);
	if (d->d_r1) hipFree(d->d_r1);
	if (d->d_order) hipFree(d->d_order);
	if (d->d_link) hipFree(d->d_link);
	if (d->stream) hipStreamDestroy(d->stream);
	free(d);
}

/* reference src/tree.c:14-55 (which memsets 0 bytes; the whole array is
   cleared here) */
int tree_init(tree_t* SF_RESTRICT tree_ptr, unsigned int max_nodes) {
	if (!tree_ptr) return STARFLOOD_FAILURE;
	tree_ptr->max_nodes = 0u;
	tree_ptr->num_nodes = 0u;
	for (int i = 0; i < 6; ++i) tree_ptr->bounds[i] = 0.0f;
	tree_ptr->node = NULL;

	const size_t mem_size = sizeof(node_t) * (size_t)max_nodes;
	void* mem = NULL;
	if (posix_memalign(&mem, (size_t)4194304u, mem_size ? mem_size : 1u) != 0 || !mem) {
		fprintf(stderr, "%s error: mem is NULL after posix_memalign(&mem, %zu, %zu)\n", "tree_init()",
		        (size_t)4194304u, mem_size);
		return STARFLOOD_FAILURE;
	}
	memset(mem, 0, mem_size);

	int slot = -1;
	for (int k = 0; k < SF_MAX_TREES; ++k)
		if (!g_tree[k]) { slot = k; break; }
	sf_tree_dev* d = slot >= 0 ? (sf_tree_dev*)calloc(1, sizeof(sf_tree_dev)) : NULL;
	if (!d) {
		sf_error("tree_init", "too many live trees");
		free(mem);
		return STARFLOOD_FAILURE;
	}
	d->key = mem;
	d->max_nodes = max_nodes;
	if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipMalloc((void**)&d->d_node, (mem_size ? mem_size : 1u)) != hipSuccess ||
	    hipMalloc((void**)&d->d_bounds, 8u * sizeof(float)) != hipSuccess ||
	    hipMalloc((void**)&d->d_meta, 4u * sizeof(unsigned int)) != hipSuccess ||
	    hipHostMalloc((void**)&d->h_meta, 4u * sizeof(unsigned int), hipHostMallocDefault) != hipSuccess) {
		sf_error("tree_init", "device allocation of %u nodes failed", max_nodes);
		tree_dev_destroy(d);
		free(mem);
		return STARFLOOD_FAILURE;
	}
	g_tree[slot] = d;
	tree_ptr->max_nodes = max_nodes;
	tree_ptr->node = (node_t*)mem;
	return STARFLOOD_SUCCESS;
}

/* reference src/tree.c:57-72 */
int tree_free(tree_t* SF_RESTRICT tree_ptr) {
	if (!tree_ptr) return STARFLOOD_FAILURE;
	if (tree_ptr->node) {
		for (int k = 0; k < SF_MAX_TREES; ++k)
			if (g_tree[k] && g_tree[k]->key == (void*)tree_ptr->node) tree_dev_destroy(g_tree[k]);
	}
	free(tree_ptr->node);
	tree_ptr->max_nodes = 0u;
	tree_ptr->num_nodes = 0u;
	for (int i = 0; i < 6; ++i) tree_ptr->bounds[i] = 0.0f;
	tree_ptr->node = NULL;
	return STARFLOOD_SUCCESS;
}

static int is_device_pointer(const void* p) {
	hipPointerAttribute_t attr;
	memset(&attr, 0, sizeof(attr));
	if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
		(void)hipGetLastError();
		return 0;
	}
	return attr.type == hipMemoryTypeDevice;
}

/* Reference-mode build from device pos/mas into d->d_node; reports the
   reference's errors (src/tree.c:180-189) after a synchronisation. */
static int build_ref(sf_tree_dev* d, tree_t* t, unsigned int N, const float* pos, const float* mas, hipStream_t s) {
	const sf_params_t* p = sf_params();
	/* 0: serial build only (also for no particles: the root alone) */
	const size_t need = N ? sfd_tree_par_bytes(N, d->max_nodes) : 0u;
	d->order_n = 0u;
	if (need && d->par_bytes < need) {
		if (d->par.scratch) {
			SF_HIP(hipStreamSynchronize(s), "tree_build");
			SF_HIP(hipStreamSynchronize(d->par.side), "tree_build");
			hipFree(d->par.scratch);
			hipFree(d->d_order);
		}
		d->par.scratch = NULL;
		d->d_order = NULL;
		d->par_bytes = 0;
		SF_HIP(hipMalloc(&d->par.scratch, need), "tree_build");
		SF_HIP(hipMalloc((void**)&d->d_order, (size_t)(N ? N : 1u) * sizeof(unsigned int)), "tree_build");
		d->par_bytes = need;
	}
	if (need && !d->par.h_state) {
		SF_HIP(hipHostMalloc((void**)&d->par.h_state, SF_TREE_PAR_STATE * sizeof(unsigned int), hipHostMallocDefault), "tree_build");
		SF_HIP(hipStreamCreateWithFlags(&d->par.side, hipStreamNonBlocking), "tree_build");
		for (int k = 0; k < 2; ++k)
			SF_HIP(hipEventCreateWithFlags(&d->par.ev[k], hipEventDisableTiming), "tree_build");
	}
	/* the level-synchronous build; a build that runs out of nodes or depth
	   is redone by the serial insertion, which stops where the reference does */
	const int rc = need ? sfd_tree_build_ref_par(&d->par, pos, mas, N, d->d_bounds, d->max_nodes,
	                                             (int)p->tree_depth_max, d->d_node, d->d_meta, d->d_order, s)
	                    : 1;
	if (rc == 0) {
		if (!d->d_link) SF_HIP(hipMalloc(&d->d_link, (size_t)d->max_nodes * 2u * sizeof(unsigned int)), "tree_build");
		if (sfd_tree_links(d->d_node, d->d_meta, d->max_nodes, d->d_link, s)) return -1;
		d->order_n = N;
	}
	if (rc < 0) return -1;
	if (rc > 0 &&
	    sfd_tree_build_ref(pos, mas, N, d->d_bounds, d->max_nodes, (int)p->tree_depth_max, d->d_node, d->d_meta, s))
		return -1;
	SF_HIP(hipMemcpyAsync(d->h_meta, d->d_meta, 2u * sizeof(unsigned int), hipMemcpyDeviceToHost, s), "tree_build");
	SF_HIP(hipStreamSynchronize(s), "tree_build");
	t->num_nodes = d->h_meta[0];
	if (p->mem_mode == SF_MEM_MIRROR && t->num_nodes) {
		SF_HIP(hipMemcpyAsync(t->node, d->d_node, (size_t)t->num_nodes * sizeof(node_t), hipMemcpyDeviceToHost, s),
		       "tree_build");
		SF_HIP(hipStreamSynchronize(s), "tree_build");
	}
	if (d->h_meta[1] == 1u) {
		fprintf(stderr, "%s error: Exceeded maximum number of tree nodes!\n", "tree_build()");
		return -1;
	}
	if (d->h_meta[1] == 2u) {
		fprintf(stderr, "%s error: Exceeded maximum tree depth!\n", "tree_build()");
		return -1;
	}
	return 0;
}

/* reference src/tree.c:74-318.  pos and mas may be host or device
   addresses; the tree's bounds are the caller's tree_ptr->bounds. */
int tree_build(tree_t* SF_RESTRICT tree_ptr, unsigned int num_parts, const real* SF_RESTRICT pos,
               const real* SF_RESTRICT mas) {
	sf_tree_dev* d = tree_dev_lookup(tree_ptr, "tree_build");
	if (!d) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	hipStream_t s = d->stream;
	const float* dpos = pos;
	const float* dmas = mas;
	if (num_parts && (!is_device_pointer(pos) || !is_device_pointer(mas))) {
		if (d->stage_n < num_parts) {
			if (d->d_stage) hipFree(d->d_stage);
			d->d_stage = NULL;
			d->stage_n = 0;
			SF_HIP(hipMalloc((void**)&d->d_stage, (size_t)num_parts * 4u * sizeof(float)), "tree_build");
			d->stage_n = num_parts;
		}
		SF_HIP(hipMemcpyAsync(d->d_stage, pos, (size_t)num_parts * 3u * sizeof(float), hipMemcpyHostToDevice, s),
		       "tree_build");
		SF_HIP(hipMemcpyAsync(d->d_stage + 3u * (size_t)num_parts, mas, (size_t)num_parts * sizeof(float),
		                      hipMemcpyHostToDevice, s),
		       "tree_build");
		dpos = d->d_stage;
		dmas = d->d_stage + 3u * (size_t)num_parts;
	}
	SF_HIP(hipMemcpyAsync(d->d_bounds, tree_ptr->bounds, 6u * sizeof(float), hipMemcpyHostToDevice, s), "tree_build");
	int rc;
	if (p->tree_mode == SF_TREE_REFERENCE) {
		rc = build_ref(d, tree_ptr, num_parts, dpos, dmas, s);
	} else {
		rc = sf_lbvh_build(&d->lbvh, num_parts, dpos, dmas, sf_params()->quadrupole, s);
		if (!rc) tree_ptr->num_nodes = num_parts ? 2u * num_parts - 1u : 0u;
		if (hipStreamSynchronize(s) != hipSuccess) rc = -1;
	}
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

/* Walk of the reference tree last built (src/gravity.c:245-663): over the
   threaded links after a complete parallel build, else the reference's own
   parent-pointer walk (which also walks a partial tree as the reference
   does). */
static int walk_ref(sf_tree_dev* d, unsigned int N, const float* pos, const float* mas, float eps2, float* acc,
                    float* pot, hipStream_t s) {
	const sf_params_t* p = sf_params();
	const int phys = p->acc_mode == SF_ACC_PHYSICAL;
	if (d->order_n == N && N)
		return sfd_tree_walk_thr(pos, mas, N, d->d_order, d->d_node, d->d_link, eps2, (float)p->theta, (float)(-p->G),
		                         phys, p->kahan_solver, p->kahan_energy, acc, pot, s);
	return sfd_tree_walk_ref(pos, mas, N, NULL, d->d_node, d->max_nodes, 2u * 8u * p->tree_nodes_max, eps2,
	                         (float)p->theta, (float)(-p->G), phys, p->kahan_solver, p->kahan_energy, acc, pot, s);
}

static void bounds_jitter(unsigned int step_number, double jitter[4]) {
	/* src/simulation.c:281-296 */
	uint32_t s[4] = {0x638FE17Bu + (uint32_t)step_number, 0xB10BDCE7u, 0x7B8ADB9Bu, 0xD8C5850Eu};
	sf_pcg4d(s);
	for (int k = 0; k < 4; ++k) jitter[k] = 2.0 * (SF_INV_U32 * (double)s[k]) - 1.0;
}

/* The tree part of src/simulation.c:245-322 on device state. */
int sf_solve_gravity_tree(sf_ctx* c, tree_t* tree, unsigned int step_number, int prev_acc) {
	sf_tree_dev* d = tree_dev_lookup(tree, "solve_gravity_part_tree");
	if (!d) return -1;
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	hipStream_t s = c->stream;
	if (N == 0) {   /* no particles: an empty tree, no forces */
		tree->num_nodes = 0u;
		return 0;
	}
	const float* pos = sf_blk(c->d_mem, N, SF_E_POS);
	const float* mas = sf_blk(c->d_mem, N, SF_E_MAS);
	float* acc = sf_blk(c->d_mem, N, SF_E_ACC);
	float* pot = sf_blk(c->d_mem, N, SF_E_POT);
	const float eps2 = (float)(p->epsilon * p->epsilon);
	if (p->tree_mode == SF_TREE_REFERENCE) {
		double jit[4];
		bounds_jitter(step_number, jit);
		if (d->r1_n < N) {
			SF_HIP(hipStreamSynchronize(s), "solve_gravity_part_tree");
			if (d->d_r1) hipFree(d->d_r1);
			d->d_r1 = NULL;
			d->r1_n = 0;
			SF_HIP(hipMalloc((void**)&d->d_r1, (size_t)N * sizeof(double)), "solve_gravity_part_tree");
			d->r1_n = N;
		}
		if (N && sfd_tree_bounds_ref(pos, N, jit, d->d_r1, d->d_bounds, s)) return -1;
		const int rc = build_ref(d, tree, N, pos, mas, s);
		SF_HIP(hipMemcpyAsync(tree->bounds, d->d_bounds, 6u * sizeof(float), hipMemcpyDeviceToHost, s),
		       "solve_gravity_part_tree");
		/* the reference walks the tree even after a failed build */
		if (walk_ref(d, N, pos, mas, eps2, acc, pot, s)) return -1;
		return rc;
	}
	if (sf_lbvh_build(&d->lbvh, N, pos, mas, p->quadrupole, s)) return -1;
	tree->num_nodes = N ? 2u * N - 1u : 0u;
	/* the relative opening test needs every particle's previous acceleration:
	   without one, a walk with the theta test alone provides it */
	if (p->mac_alpha > 0.0 && !prev_acc && sf_lbvh_walk(d->lbvh, N, pos, mas, p, 0, acc, pot, s)) return -1;
	return sf_lbvh_walk(d->lbvh, N, pos, mas, p, p->mac_alpha > 0.0, acc, pot, s);
}

int sf_lbvh_export(const tree_t* tree_ptr, unsigned long long* keys, unsigned int* order,
                   unsigned long long counts[4], float* node_cm, int* node_lr_range, float* node_cell) {
	sf_tree_dev* d = tree_dev_lookup(tree_ptr, "sf_lbvh_export");
	if (!d) return STARFLOOD_FAILURE;
	SF_HIP(hipDeviceSynchronize(), "sf_lbvh_export");
	return sf_lbvh_export_impl(d->lbvh, d->stream, keys, order, counts, node_cm, node_lr_range, node_cell)
	           ? STARFLOOD_FAILURE
	           : STARFLOOD_SUCCESS;
}

/* After LBVH walks on stream s: fail if any wavefront of the walks since
   the last check met a malformed tree or ran out of stack (synchronises s). */
static int lbvh_check(sf_tree_dev* d, hipStream_t s, const char* fn) {
	unsigned long long tripped = 0;
	if (sf_lbvh_guard(d->lbvh, s, &tripped)) return -1;
	if (tripped) {
		sf_error(fn, "%llu wavefronts of the tree walk met a malformed tree: forces are incomplete", tripped);
		return -1;
	}
	return 0;
}

int sf_tree_check(tree_t* tree, hipStream_t s, const char* fn) {
	if (sf_params()->tree_mode != SF_TREE_LBVH) return 0;
	sf_tree_dev* d = tree_dev_lookup(tree, fn);
	return d ? lbvh_check(d, s, fn) : -1;
}

/* reference src/gravity.c:245-663: walks the tree last built for this
   simulation (tree_build or sim_solv) and writes acc and pot. */
static int gravity_part_tree(const sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "solve_gravity_part_tree");
	if (!c) return STARFLOOD_FAILURE;
	sf_tree_dev* d = tree_dev_lookup(&sim_ptr->tree, "solve_gravity_part_tree");
	if (!d) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	hipStream_t s = c->stream;
	if (sf_upload_blocks(c, SF_E_POS, 3u) || sf_upload_blocks(c, SF_E_MAS, 1u)) return STARFLOOD_FAILURE;
	/* order the tree's stream before the simulation's */
	SF_HIP(hipStreamSynchronize(d->stream), "solve_gravity_part_tree");
	const float* pos = sf_blk(c->d_mem, N, SF_E_POS);
	const float* mas = sf_blk(c->d_mem, N, SF_E_MAS);
	float* acc = sf_blk(c->d_mem, N, SF_E_ACC);
	float* pot = sf_blk(c->d_mem, N, SF_E_POT);
	int rc;
	if (p->tree_mode == SF_TREE_REFERENCE) {
		rc = walk_ref(d, N, pos, mas, (float)(p->epsilon * p->epsilon), acc, pot, s);
	} else {
		/* standalone: acc holds no previous force, so the relative opening
		   test (mac_alpha > 0) starts from a theta-only walk */
		rc = d->lbvh ? sf_lbvh_walk(d->lbvh, N, pos, mas, p, 0, acc, pot, s) : -1;
		if (!rc && d->lbvh && p->mac_alpha > 0.0) rc = sf_lbvh_walk(d->lbvh, N, pos, mas, p, 1, acc, pot, s);
		if (!d->lbvh) sf_error("solve_gravity_part_tree", "no tree was built");
		if (!rc && lbvh_check(d, s, "solve_gravity_part_tree")) rc = -1;
	}
	if (sf_download_blocks(c, SF_E_ACC, 3u) || sf_download_blocks(c, SF_E_POT, 1u)) rc = -1;
	if (sf_finish(c, "solve_gravity_part_tree")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

/* the prototype form the caller's SOLVER_VIS selects (reference
   src/gravity.h:14-18); the debug overlay is not drawn */
#ifdef SOLVER_VIS
int solve_gravity_part_tree(const sim_t* SF_RESTRICT sim_ptr, f32* render_buffer, unsigned int sizex,
                            unsigned int sizey) {
	(void)render_buffer;
	(void)sizex;
	(void)sizey;
	return gravity_part_tree(sim_ptr);
}
#else
int solve_gravity_part_tree(const sim_t* SF_RESTRICT sim_ptr) { return gravity_part_tree(sim_ptr); }
#endif
