/*
 * tree_dev.h -- device-side tree records and launchers shared by the tree
 * kernels and host/tree.c.
 */
#ifndef SF_TREE_DEV_H
#define SF_TREE_DEV_H

#include "../sf_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same 80-byte layout as node_t (reference src/tree.h:16-25), so the device
   node array of the reference build copies straight into tree_t.node. */
typedef struct {
	unsigned int parent;
	unsigned int bodies;
	unsigned int branch[8];
	float bounds[6];
	float params[4];
} sf_node;

/* --- reference octree (tree_ref.hip) --- */
int sfd_tree_bounds_ref(const float* pos, unsigned int N, const double jitter[3], float* bounds, hipStream_t s);
/* meta[0] = num_nodes, meta[1] = status (0 ok, 1 out of nodes, 2 too deep) */
int sfd_tree_build_ref(const float* pos, const float* mas, unsigned int N, const float* bounds, unsigned int max_nodes,
                       int depth_max, sf_node* node, unsigned int* meta, hipStream_t s);
int sfd_tree_walk_ref(const float* pos, const float* mas, unsigned int N, const sf_node* node, unsigned int node_cap,
                      unsigned int iter_max,
                      float eps2, float theta, float negG, int acc_phys, int kahan_acc, int kahan_pot, float* acc,
                      float* pot, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
