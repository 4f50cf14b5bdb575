/*
 * lbvh.c -- host side of the SF_TREE_LBVH Barnes-Hut tree.
 */
#include <stdlib.h>

#include "sf_host.h"

struct sf_lbvh {
	unsigned int n_alloc;
};

int sf_lbvh_build(sf_lbvh** st, unsigned int N, const float* pos, const float* mas, hipStream_t s) {
	(void)st; (void)N; (void)pos; (void)mas; (void)s;
	sf_error("tree_build", "SF_TREE_LBVH is not built in this library version");
	return -1;
}

int sf_lbvh_walk(sf_lbvh* st, unsigned int N, const float* pos, const float* mas, const sf_params_t* p, float* acc,
                 float* pot, hipStream_t s) {
	(void)st; (void)N; (void)pos; (void)mas; (void)p; (void)acc; (void)pot; (void)s;
	sf_error("solve_gravity_part_tree", "SF_TREE_LBVH is not built in this library version");
	return -1;
}

void sf_lbvh_destroy(sf_lbvh* st) { free(st); }
