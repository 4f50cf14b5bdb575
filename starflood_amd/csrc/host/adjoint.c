/*
 * adjoint.c -- forward KDK with checkpoints and its adjoint (no reference
 * counterpart; see include/starflood_ext.h).
 */
#include <stdlib.h>

#include "sf_host.h"

struct sf_tape {
	unsigned int N;
};

int sf_adjoint_forward(unsigned int N, const float* x0, const float* v0, const float* m, unsigned int nsteps,
                       unsigned int ckpt_every, float* xN, float* vN, void* stream, sf_tape** tape_out) {
	(void)N; (void)x0; (void)v0; (void)m; (void)nsteps; (void)ckpt_every; (void)xN; (void)vN; (void)stream;
	if (tape_out) *tape_out = NULL;
	sf_error("sf_adjoint_forward", "not built in this library version");
	return STARFLOOD_FAILURE;
}

int sf_adjoint_backward(sf_tape* tape, const float* gxN, const float* gvN, float* gx0, float* gv0, void* stream) {
	(void)tape; (void)gxN; (void)gvN; (void)gx0; (void)gv0; (void)stream;
	sf_error("sf_adjoint_backward", "not built in this library version");
	return STARFLOOD_FAILURE;
}

void sf_tape_free(sf_tape* tape) { free(tape); }

int sf_force_dev(unsigned int N, const float* x, const float* m, float* acc, float* pot, void* stream) {
	(void)N; (void)x; (void)m; (void)acc; (void)pot; (void)stream;
	sf_error("sf_force_dev", "not built in this library version");
	return STARFLOOD_FAILURE;
}

int sf_force_vjp_dev(unsigned int N, const float* x, const float* m, const float* mu, float* g, void* stream) {
	(void)N; (void)x; (void)m; (void)mu; (void)g; (void)stream;
	sf_error("sf_force_vjp_dev", "not built in this library version");
	return STARFLOOD_FAILURE;
}
