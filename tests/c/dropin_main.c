/* A Starflood driver in the shape of the reference program's main()
 * (src/main.c:348-611), linked against libstarflood_hip.so in place of the
 * reference's gravity.c, sph.c, tree.c, simulation.c, initcond.c and
 * simulation_io.c.  It includes the library's headers, sets the run-time
 * parameters that stand in for config.h's macros once before sim_init, and
 * then makes the reference program's calls in its order:
 *   sim_init (src/main.c:358), initcond_generate (:384), sim_solv (:398/400),
 *   per step: stop file (:427-446), step_number (:451), sim_save_raw before
 *   the step (:486-495), sim_step (:543/545); sim_free (:597).
 * Compile-time settings, as config.h gives the reference program:
 *   NUM_BODIES, NUM_TIMESTEPS, OUTPUT_INTERVAL (0: every step), EPSILON,
 *   TIMESTEP_SIZE, OUTPUT_DIR, SIM_FILENAME; SOLVER_VIS selects the
 *   render-buffer prototypes (link libstarflood_hip_solvervis.so then).
 * Exit status: EXIT_SUCCESS, or EXIT_FAILURE when set-up or a step fails. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "starflood.h"
#include "starflood_ext.h"

#ifndef NUM_BODIES
#define NUM_BODIES 256u
#endif
#ifndef NUM_TIMESTEPS
#define NUM_TIMESTEPS 10u
#endif
#ifndef OUTPUT_INTERVAL
#define OUTPUT_INTERVAL 0u
#endif
#ifndef EPSILON
#define EPSILON 1.0e-2
#endif
#ifndef TIMESTEP_SIZE
#define TIMESTEP_SIZE 1.0e-3
#endif
#ifndef OUTPUT_DIR
#define OUTPUT_DIR "./out"
#endif
#ifndef SIM_FILENAME
#define SIM_FILENAME OUTPUT_DIR "/sim/step_%04u.raw"
#endif

#ifdef SOLVER_VIS
#define SOLV(s) sim_solv((s), NULL, 0u, 0u)
#define STEP(s) sim_step((s), NULL, 0u, 0u)
#else
#define SOLV(s) sim_solv(s)
#define STEP(s) sim_step(s)
#endif

static int stop_requested(void) {
	FILE* f = fopen(OUTPUT_DIR "/stop", "r");
	if (f == NULL) return 0;
	fclose(f);
	if (remove(OUTPUT_DIR "/stop") != 0) perror("remove(" OUTPUT_DIR "/stop)");
	return 1;
}

int main(void) {
	sf_params_t p;
	sf_params_default(&p);
	p.epsilon = EPSILON;
	p.dt = TIMESTEP_SIZE;
	p.enable_tree = 0;
	strncpy(p.out_dir, OUTPUT_DIR, sizeof p.out_dir - 1u);
	p.out_dir[sizeof p.out_dir - 1u] = '\0';
	if (sf_set_params(&p) != STARFLOOD_SUCCESS) {
		fprintf(stderr, "fatal error: sf_set_params() failed.\n");
		return EXIT_FAILURE;
	}

	sim_t sim;
	if (sim_init(&sim, NUM_BODIES) != STARFLOOD_SUCCESS) {
		fprintf(stderr, "fatal error: sim_init(&sim, %u) failed.\n", (unsigned)NUM_BODIES);
		return EXIT_FAILURE;
	}
	if (initcond_generate(&sim) != STARFLOOD_SUCCESS || SOLV(&sim) != STARFLOOD_SUCCESS) {
		fprintf(stderr, "fatal error: initial conditions or the initial solve failed.\n");
		sim_free(&sim);
		return EXIT_FAILURE;
	}

	const unsigned interval = OUTPUT_INTERVAL > 0u ? (unsigned)OUTPUT_INTERVAL : 1u;
	int rc = EXIT_SUCCESS;
	char name[512];
	for (unsigned step = 0u; step <= (unsigned)NUM_TIMESTEPS; step++) {
		if (stop_requested()) break;
		sim.step_number = step;
		if (step % interval == 0u) {
			snprintf(name, sizeof name, SIM_FILENAME, step / interval);
			if (sim_save_raw(&sim, name) != STARFLOOD_SUCCESS)
				fprintf(stderr, "error: sim_save_raw(&sim, \"%s\") failed.\n", name);
		}
		if (step < (unsigned)NUM_TIMESTEPS && STEP(&sim) != STARFLOOD_SUCCESS) {
			fprintf(stderr, "error: sim_step(&sim) failed.\n");
			rc = EXIT_FAILURE;
			break;
		}
	}
	if (sim_free(&sim) != STARFLOOD_SUCCESS) {
		fprintf(stderr, "error: sim_free(&sim) failed.\n");
		rc = EXIT_FAILURE;
	}
	return rc;
}
