/*
 * driver.c -- the reference program's run loop (src/main.c:348-611) as a
 * library call, over this library's sim_* and vis_* entry points.
 *
 * Order of one step, as src/main.c:423-591:
 *   stop file <out_dir>/stop present -> remove it and end the run;
 *   sim.step_number = step;
 *   on output steps: save the snapshot (before the step), or read it when
 *   the simulation is disabled;
 *   on output steps: vis_draw;
 *   step < num_timesteps: sim_step;
 *   on output steps: un-premultiply alpha and vis_save.
 * Steps run from 0 to num_timesteps inclusive; with the simulation disabled
 * and an output interval, the step advances by the interval.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sf_host.h"
#include "../../../include/starflood_vis.h"

/* A file-name pattern may hold literal text, "%%", and exactly one
   conversion of an unsigned int: '%', optional '0' / '-' flags, an optional
   width, 'u'.  Anything else is refused, so a pattern can never read
   arguments that are not there. */
static int pattern_ok(const char* p) {
	int conv = 0;
	for (; *p; ++p) {
		if (*p != '%') continue;
		++p;
		if (*p == '%') continue;
		while (*p == '0' || *p == '-') ++p;
		while (*p >= '0' && *p <= '9') ++p;
		if (*p != 'u') return 0;
		++conv;
	}
	return conv == 1;
}

/* <dir>/<name><ext> as a pattern: a '%' of the directory is doubled, so
   only the name's conversion formats */
static int make_pattern(char* out, size_t cap, const char* dir, const char* name, const char* ext) {
	size_t n = 0;
	for (const char* c = dir; *c; ++c) {
		if (n + 2u >= cap) return 0;
		if (*c == '%') out[n++] = '%';
		out[n++] = *c;
	}
	const int m = snprintf(out + n, cap - n, "/%s%s", name, ext);
	return m > 0 && (size_t)m < cap - n;
}

/* snprintf with a pattern checked by pattern_ok */
static int format_name(char* out, size_t cap, const char* pattern, unsigned int num) {
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wformat-nonliteral"
	const int n = snprintf(out, cap, pattern, num);
#pragma GCC diagnostic pop
	return n > 0 && (size_t)n < cap;
}

static const char* vis_ext(image_filetype_t t) {
	switch (t) {
	case IMAGE_FILETYPE_PFM: return ".pfm";
	case IMAGE_FILETYPE_PPM: return ".ppm";
	default: return ".png";
	}
}

/* src/main.c:558-582: divide RGB by alpha (when non-zero), alpha = 1 */
static void unpremultiply(vis_t* v) {
	const size_t n = (size_t)v->sizex * v->sizey;
	float* px = v->render_buffer;
	for (size_t i = 0; i < n; ++i) {
		float* c = px + 4u * i;
		const float inv_alpha = 0.0f != c[3] ? 1.0f / c[3] : 1.0f;
		c[0] *= inv_alpha;
		c[1] *= inv_alpha;
		c[2] *= inv_alpha;
		c[3] = 1.0f;
	}
}

int sf_run(const sf_run_config_t* cfg) {
	static const char* fn = "sf_run";
	if (!cfg) {
		sf_error(fn, "cfg is NULL");
		return STARFLOOD_FAILURE;
	}
	const int enable_sim = cfg->enable_sim != 0, enable_vis = cfg->enable_vis != 0;
	const int sim_io = cfg->sim_filename != NULL, vis_io = cfg->vis_filename != NULL;
	const unsigned int oi = cfg->output_interval;
	if ((unsigned int)cfg->vis_file_type >= (unsigned int)NUM_IMAGE_FILETYPE) {
		sf_error(fn, "unknown vis_file_type %d", (int)cfg->vis_file_type);
		return STARFLOOD_FAILURE;
	}
	sf_params_t p;
	sf_get_params(&p);
	enum { CAP = 4096 };
	char sim_pattern[CAP], vis_pattern[CAP], sim_name[CAP], vis_name[CAP], stop_name[CAP];
	if (snprintf(stop_name, CAP, "%s/stop", p.out_dir) >= (int)CAP ||
	    (sim_io && !make_pattern(sim_pattern, CAP, p.out_dir, cfg->sim_filename, ".raw")) ||
	    (vis_io && !make_pattern(vis_pattern, CAP, p.out_dir, cfg->vis_filename, vis_ext(cfg->vis_file_type)))) {
		sf_error(fn, "output path too long");
		return STARFLOOD_FAILURE;
	}
	if ((sim_io && !pattern_ok(sim_pattern)) || (vis_io && !pattern_ok(vis_pattern))) {
		sf_error(fn, "a file name pattern needs exactly one %%u conversion (flags 0 and -, a width)");
		return STARFLOOD_FAILURE;
	}
	if (enable_vis && (cfg->vis_sizex == 0u || cfg->vis_sizey == 0u)) {
		sf_error(fn, "enable_vis needs vis_sizex and vis_sizey");
		return STARFLOOD_FAILURE;
	}
	/* the cubic model's u lives on the device only: the run loop has no
	   initial value for it and raw snapshots do not carry it */
	if (p.enable_sph && p.sph_model == SF_SPH_CUBIC_IDEAL) {
		sf_error(fn, "the run loop does not carry the cubic SPH model's internal energy u (not in raw "
		             "snapshots); use sph_model SF_SPH_REFERENCE, or drive sim_step directly with "
		             "sf_ext_copy(SF_EXT_U)");
		return STARFLOOD_FAILURE;
	}

	/* src/main.c:348-412: visualisation, simulation, initial conditions */
	vis_t vis;
	sim_t sim;
	memset(&vis, 0, sizeof(vis));
	memset(&sim, 0, sizeof(sim));
	if (enable_vis && STARFLOOD_SUCCESS != vis_init(&vis, cfg->vis_sizex, cfg->vis_sizey)) {
		sf_error(fn, "vis_init(%u, %u) failed", cfg->vis_sizex, cfg->vis_sizey);
		return STARFLOOD_FAILURE;
	}
	int rc = STARFLOOD_SUCCESS;
	if (STARFLOOD_SUCCESS != sim_init(&sim, cfg->num_bodies)) {
		sf_error(fn, "sim_init(%u) failed", cfg->num_bodies);
		rc = STARFLOOD_FAILURE;
		goto free_vis;
	}
	if (cfg->init_cond_file) {
		/* src/main.c:370-381: acc comes from the snapshot, no initial solve */
		if (STARFLOOD_SUCCESS != sim_read_raw(&sim, cfg->init_cond_file)) {
			sf_error(fn, "sim_read_raw(\"%s\") failed", cfg->init_cond_file);
			rc = STARFLOOD_FAILURE;
			goto free_sim;
		}
	} else if (enable_sim) {
		if (STARFLOOD_SUCCESS != initcond_generate(&sim)) {
			sf_error(fn, "initcond_generate() failed");
			rc = STARFLOOD_FAILURE;
			goto free_sim;
		}
		if (STARFLOOD_SUCCESS != sf_sim_solv(&sim)) {
			sf_error(fn, "sim_solv() failed");
			rc = STARFLOOD_FAILURE;
			goto free_sim;
		}
	}

	/* src/main.c:414-421 */
	const unsigned int step_inc = (!enable_sim && oi) ? oi : 1u;
	for (unsigned int step = 0u; step <= cfg->num_timesteps; step += step_inc) {
		if (cfg->verbose) {
			printf("Step #%u (run is %6.03f%% complete)\n", step,
			       100.0 * (0u < cfg->num_timesteps ? (double)step / (double)cfg->num_timesteps : 1.0));
			printf("t = %.06f\n", p.dt * (double)step);
		}
		/* src/main.c:427-446: the stop file ends the run and is removed */
		FILE* stopfile = fopen(stop_name, "r");
		if (stopfile) {
			if (cfg->verbose) printf("\nStopfile found! Stopping run...\n");
			fclose(stopfile);
			if (0 != remove(stop_name)) sf_error(fn, "remove(\"%s\") failed", stop_name);
			break;
		}
		fflush(stdout);
		sim.step_number = step;
		const int output_step = oi ? (0u == step % oi) : 1;
		const unsigned int file_num = oi ? step / oi : step;
		if ((sim_io && !format_name(sim_name, CAP, sim_pattern, file_num)) ||
		    (vis_io && !format_name(vis_name, CAP, vis_pattern, file_num))) {
			sf_error(fn, "file name for step %u too long", step);
			rc = 1;
			break;
		}
		/* src/main.c:487-528: snapshot before the step, or replay */
		if (output_step && sim_io) {
			if (enable_sim) {
				if (STARFLOOD_SUCCESS != sim_save_raw(&sim, sim_name))
					fprintf(stderr, "error: sim_save_raw(&sim, \"%s\") failed.\n", sim_name);
			} else if (STARFLOOD_SUCCESS != sim_read_raw(&sim, sim_name)) {
				fprintf(stderr, "fatal error: sim_read_raw(&sim, \"%s\") failed.\n", sim_name);
				rc = 1;
				break;
			}
		}
		/* src/main.c:530-539 */
		if (output_step && enable_vis && STARFLOOD_SUCCESS != vis_draw(&vis, &sim))
			fprintf(stderr, "error: vis_draw(&vis, &sim) failed.\n");
		/* src/main.c:541-550 */
		if (step < cfg->num_timesteps && enable_sim && STARFLOOD_SUCCESS != sf_sim_step(&sim))
			fprintf(stderr, "error: sim_step(&sim) failed.\n");
		/* src/main.c:552-588 */
		if (output_step && enable_vis && vis_io) {
			unpremultiply(&vis);
			if (STARFLOOD_SUCCESS != vis_save(&vis, vis_name, cfg->vis_file_type))
				fprintf(stderr, "error: vis_save(&vis, \"%s\") failed.\n", vis_name);
		}
		if (cfg->verbose) printf("\n");
		/* the next step exceeds num_timesteps (also where step + step_inc would wrap) */
		if (cfg->num_timesteps >= step_inc && step > cfg->num_timesteps - step_inc) break;
	}

free_sim:
	if (STARFLOOD_SUCCESS != sim_free(&sim)) {
		fprintf(stderr, "error: sim_free(&sim) failed.\n");
		if (rc == STARFLOOD_SUCCESS) rc = STARFLOOD_FAILURE;
	}
free_vis:
	if (enable_vis && STARFLOOD_SUCCESS != vis_free(&vis)) {
		fprintf(stderr, "error: vis_free(&vis) failed.\n");
		if (rc == STARFLOOD_SUCCESS) rc = STARFLOOD_FAILURE;
	}
	return rc;
}
