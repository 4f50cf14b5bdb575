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
/* the reference's FP64 Kahan mean radius into avg[0] (r1: N doubles of scratch) */
int sfd_mean_radius(const float* pos, unsigned int N, double* r1, double* avg, hipStream_t s);
/* r1: N doubles of scratch */
int sfd_tree_bounds_ref(const float* pos, unsigned int N, const double jitter[3], double* r1, float* bounds,
                        hipStream_t s);
/* meta[0] = num_nodes, meta[1] = status (0 ok, 1 out of nodes, 2 too deep) */
int sfd_tree_build_ref(const float* pos, const float* mas, unsigned int N, const float* bounds, unsigned int max_nodes,
                       int depth_max, sf_node* node, unsigned int* meta, hipStream_t s);
/* Level-synchronous build of the same tree (tree_ref_par.hip).  scratch:
   sfd_tree_par_bytes(N, max_nodes) bytes of device memory (0: only the
   serial build applies, N > 8 max_nodes); h_state: SF_TREE_PAR_STATE u32 of pinned host
   memory; side: a second stream and ev: two events for the root's moment
   sums.  order (N u32) receives a walk order: the particles leaf by leaf,
   those outside the bounds last.  Returns 0 when node and meta hold the
   serial build's result, 1 when
   the build would exceed max_nodes or depth_max (the caller then runs
   sfd_tree_build_ref, which leaves the partial tree the reference leaves),
   -1 on a HIP error. */
#define SF_TREE_PAR_STATE 16
typedef struct {
	void* scratch;
	unsigned int* h_state;
	hipStream_t side;
	hipEvent_t ev[2];
} sf_tree_par_ctx;
size_t sfd_tree_par_bytes(unsigned int N, unsigned int max_nodes);
int sfd_tree_build_ref_par(const sf_tree_par_ctx* x, const float* pos, const float* mas, unsigned int N,
                           const float* bounds, unsigned int max_nodes, int depth_max, sf_node* node,
                           unsigned int* meta, unsigned int* order, hipStream_t s);
/* Threaded links of a complete reference tree (meta[0] nodes; link: max_nodes
   x 2 u32): each node's first non-empty child and the first non-empty node
   after its subtree in the walk's order.  sfd_tree_walk_thr walks them and
   adds the same terms in the same order as sfd_tree_walk_ref. */
int sfd_tree_links(const sf_node* node, const unsigned int* meta, unsigned int max_nodes, void* link, hipStream_t s);
int sfd_tree_walk_thr(const float* pos, const float* mas, unsigned int N, const unsigned int* order,
                      const sf_node* node, const void* link, float eps2, float theta, float negG, int acc_phys,
                      int kahan_acc, int kahan_pot, float* acc, float* pot, hipStream_t s);
/* order: the particle of each walk lane (N entries), or NULL for index order */
int sfd_tree_walk_ref(const float* pos, const float* mas, unsigned int N, const unsigned int* order, const sf_node* node,
                      unsigned int node_cap, unsigned int iter_max,
                      float eps2, float theta, float negG, int acc_phys, int kahan_acc, int kahan_pot, float* acc,
                      float* pot, hipStream_t s);

/* --- LBVH tree (lbvh.hip); node arrays are float4 / int2 / int4 on the
   device, passed untyped through C --- */
unsigned int sfd_lbvh_bbox_blocks(void);
unsigned int sfd_radix_tile(void);
/* u32 entries of the radix-sort histogram scratch for n_sort keys */
unsigned int sfd_radix_hist_entries(unsigned int n_sort);
/* bbox (FP64), then 63-bit Morton keys and identity values for n_sort
   (a multiple of the radix tile) entries; padding keys are all ones.
   packed: N float4 (x, y, z, m) in index order, for sfd_lbvh_topology */
int sfd_lbvh_keys(const float* pos, const float* mas, unsigned int N, unsigned int n_sort, double* bbox_part,
                  double* box, unsigned long long* key, unsigned int* val, void* packed, hipStream_t s);
/* stable 8-pass LSD radix sort; the result is in (key0, val0) */
int sfd_radix_sort(unsigned long long* key0, unsigned int* val0, unsigned long long* key1, unsigned int* val1,
                   unsigned int n_sort, unsigned int* hist, hipStream_t s);
/* the first `passes` 8-bit passes only: the pairs are sorted on the low
   8 * passes key bits and end in (key0, val0) for an even pass count,
   (key1, val1) for an odd one */
int sfd_radix_sort_passes(unsigned long long* key0, unsigned int* val0, unsigned long long* key1, unsigned int* val1,
                          unsigned int n_sort, unsigned int* hist, unsigned int passes, hipStream_t s);
/* in-place exclusive scan of n u32 entries with sfd_scan_u32_sums_len(n)
   entries of scratch */
unsigned int sfd_scan_u32_sums_len(unsigned int n);
int sfd_scan_u32_exclusive(unsigned int* data, unsigned int n, unsigned int* sums, hipStream_t s);
/* tiles of the FP64 moment scan (records of 4 doubles each) */
unsigned int sfd_lbvh_scan_tiles(unsigned int N);
/* sorted bodies, radix tree, FP64 prefix moments (N records of 4 doubles),
   node monopoles (2N-1 float4) and Morton cells (2N-1 float4: centre, side;
   internal nodes of at most 64 particles are left to sfd_lbvh_group_boxes).
   prefix holds sfd_lbvh_keys' packed records on entry */
int sfd_lbvh_topology(const unsigned long long* key, const unsigned int* order, unsigned int N, const double* box,
                      void* body, void* node_lr, void* node_range, void* prefix, void* tile_sum, void* node_cm,
                      void* node_cell, hipStream_t s);
int sfd_lbvh_records(unsigned int N, const void* node_lr, const void* node_range, const void* node_cm,
                     const void* node_cell, float theta, int leaf_size, void* rec_b, hipStream_t s);
/* lbvh_walk.hip.  counters (6 x u64; all but the fifth zeroed here): cell
   interactions, particle interactions, wavefronts that met a malformed tree,
   node batches; the fifth adds the third up over walks until the host clears
   it; the sixth hands out the chunks of 64 sorted positions.
   gstack: sfd_lbvh_walk_stack_bytes() of scratch.  alpha_g > 0 adds the
   relative opening test M l^p <= alpha_g |a|_min d^(p+2) (l the records'
   size, p = 2 for monopole cells and 3 for quadrupole cells; a quadrupole
   walk takes the cells that meet p = 2 as monopoles), |a|_min over each
   group's previous acc. */
size_t sfd_lbvh_walk_stack_bytes(void);
int sfd_lbvh_walk(const void* body, const unsigned int* order, unsigned int N, const unsigned int* group_start,
                  const void* node_cm, const void* node_q0, const float* node_q1, const void* rec_b,
                  float alpha_g, float eps2, double G, int acc_phys, float* acc, float* pot,
                  unsigned long long* counters, int* gstack, hipStream_t s);
/* traceless quadrupoles (node_q0 2N-1 float4, node_q1 2N-1 float) from a
   second FP64 prefix scan (prefix2: N x 6 doubles, tile_sum2:
   sfd_lbvh_scan_tiles(N) x 6 doubles, ref: 3 doubles); NULL node_q0 in
   sfd_lbvh_walk selects the monopole walk */
int sfd_lbvh_quadrupoles(unsigned int N, const void* body, const void* node_range, const void* prefix,
                         const double* box, void* prefix2, void* tile_sum2, double* ref, void* node_q0,
                         float* node_q1, hipStream_t s);

/* walk groups: start[p] = 1 where a maximal subtree of at most cap (<= 64)
   particles begins in sorted order, and start[N] = 1; start holds
   sfd_lbvh_group_flag_len(N) entries (lbvh.hip) */
unsigned int sfd_lbvh_group_flag_len(unsigned int N);
int sfd_lbvh_groups(unsigned int N, const void* node_lr, const void* node_range, int cap, unsigned int* start,
                    hipStream_t s);
/* the particle boxes of the internal nodes of at most 64 particles into
   node_cell (their other entries come from sfd_lbvh_topology); start: the
   groups of sfd_lbvh_groups with cap 64 */
int sfd_lbvh_group_boxes(unsigned int N, const void* body, const unsigned int* start, const void* node_range,
                         void* node_cell, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif
