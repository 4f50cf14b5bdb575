/*
 * sph.c -- solve_sph with the reference's signature (src/sph.h:6,
 * src/sph.c:14-283): density, equation of state, pressure force added into
 * acc.
 */
#include <stdio.h>

#include "sf_host.h"

int sf_solve_sph(sf_ctx* c) {
	(void)c;
	sf_error("solve_sph", "SPH kernels are not built in this library version");
	return -1;
}

int solve_sph(const sim_t* SF_RESTRICT sim_ptr) {
	if (!sim_ptr) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "solve_sph");
	if (!c) return STARFLOOD_FAILURE;
	if (!sf_params()->enable_sph) {
		sf_error("solve_sph", "enable_sph is 0 (sim.mem has no rad/rho/prs blocks)");
		return STARFLOOD_FAILURE;
	}
	if (sf_upload_blocks(c, SF_E_POS, 4u) || sf_upload_blocks(c, SF_E_ACC, 4u) || sf_upload_blocks(c, SF_E_RAD, 1u))
		return STARFLOOD_FAILURE;
	int rc = sf_solve_sph(c);
	if (sf_download_blocks(c, SF_E_ACC, 3u) || sf_download_blocks(c, SF_E_RHO, 2u)) rc = -1;
	if (sf_finish(c, "solve_sph")) rc = -1;
	return rc ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
