/*
 * params.c -- run-time parameters (the reference's compile-time src/config.h
 * knobs), error reporting and version.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../sf_internal.h"

#define SF_VERSION "starflood-mi355x 0.1.0 (gfx950)"

static sf_params_t g_params;
static int g_params_ready = 0;
static _Thread_local char g_err[512];

/* defaults are the reference's src/config.h values */
void sf_params_default(sf_params_t* p) {
	memset(p, 0, sizeof(*p));
	p->epsilon = 1.000e-3;          /* config.h:173 */
	p->G = 1.000;                   /* config.h:178 */
	p->dt = 1.000e-0;               /* config.h:170 */
	p->theta = 0.5;                 /* config.h:44 */
	p->tree_nodes_max = 262144u;    /* config.h:41 */
	p->tree_depth_max = 16u;        /* config.h:38 */
	p->leaf_size = 3u;               /* fastest measured at theta 0.5 (2..6, 8, 16), same error */
	p->enable_tree = 1;             /* config.h:35 */
	p->enable_sph = 0;              /* config.h:190 */
	p->kahan_solver = 1;            /* config.h:194 */
	p->kahan_energy = 1;            /* config.h:198 */
	p->precision = SF_PRECISION_EXACT;
	p->acc_mode = SF_ACC_REFERENCE;
	p->tree_mode = SF_TREE_REFERENCE;
	p->sph_model = SF_SPH_REFERENCE;
	p->mem_mode = SF_MEM_MIRROR;
	p->sph_gamma = 5.0 / 3.0;
	p->logging = 1;                 /* config.h:120-123 */
	p->verbose = 0;
	strcpy(p->out_dir, "./out");    /* config.h:75 */
	p->quadrupole = 0;
	p->sph_search = SF_SPH_SEARCH_GRID;
	p->mac_alpha = 0.0;
}

const sf_params_t* sf_params(void) {
	if (!g_params_ready) {
		sf_params_default(&g_params);
		g_params_ready = 1;
	}
	return &g_params;
}

/* context.c: is a simulation alive whose sim.mem layout differs from dof? */
int sf_ctx_dof_conflict(unsigned int dof);

int sf_set_params(const sf_params_t* p) {
	if (!p) {
		sf_error("sf_set_params", "params is NULL");
		return STARFLOOD_FAILURE;
	}
	if (!(p->epsilon >= 0.0) || !(p->dt == p->dt) || !(p->theta >= 0.0)) {
		sf_error("sf_set_params", "epsilon and theta must be >= 0 and dt a number");
		return STARFLOOD_FAILURE;
	}
	if (!(p->mac_alpha >= 0.0) || p->mac_alpha > 1.0e30) {
		sf_error("sf_set_params", "mac_alpha must be finite and >= 0");
		return STARFLOOD_FAILURE;
	}
	if (p->tree_depth_max < 1u || p->tree_depth_max > 16u) {
		sf_error("sf_set_params", "tree_depth_max %u outside [1, 16]", p->tree_depth_max);
		return STARFLOOD_FAILURE;
	}
	if (p->leaf_size < 1u || p->leaf_size > 64u) {
		sf_error("sf_set_params", "leaf_size %u outside [1, 64]", p->leaf_size);
		return STARFLOOD_FAILURE;
	}
	if ((p->precision != SF_PRECISION_EXACT && p->precision != SF_PRECISION_FAST) ||
	    (p->acc_mode != SF_ACC_REFERENCE && p->acc_mode != SF_ACC_PHYSICAL) ||
	    (p->tree_mode != SF_TREE_REFERENCE && p->tree_mode != SF_TREE_LBVH) ||
	    (p->sph_model != SF_SPH_REFERENCE && p->sph_model != SF_SPH_CUBIC_IDEAL) ||
	    (p->sph_search != SF_SPH_SEARCH_ALL && p->sph_search != SF_SPH_SEARCH_GRID) ||
	    (p->mem_mode != SF_MEM_MIRROR && p->mem_mode != SF_MEM_RESIDENT)) {
		sf_error("sf_set_params", "unknown enumeration value");
		return STARFLOOD_FAILURE;
	}
	if (p->quadrupole != 0 && p->quadrupole != 1) {
		sf_error("sf_set_params", "quadrupole must be 0 or 1");
		return STARFLOOD_FAILURE;
	}
	if (memchr(p->out_dir, 0, sizeof(p->out_dir)) == NULL) {
		sf_error("sf_set_params", "out_dir is not NUL-terminated");
		return STARFLOOD_FAILURE;
	}
	const unsigned int dof = p->enable_sph ? 16u : 13u;
	if (sf_ctx_dof_conflict(dof)) {
		sf_error("sf_set_params", "enable_sph changes the sim.mem layout of a live simulation");
		return STARFLOOD_FAILURE;
	}
	g_params = *p;
	g_params_ready = 1;
	return STARFLOOD_SUCCESS;
}

int sf_get_params(sf_params_t* p) {
	if (!p) return STARFLOOD_FAILURE;
	*p = *sf_params();
	return STARFLOOD_SUCCESS;
}

unsigned int sf_sim_dof(void) { return sf_params()->enable_sph ? 16u : 13u; }

const char* sf_last_error(void) { return g_err; }

const char* sf_version(void) { return SF_VERSION; }

void sf_error(const char* func, const char* fmt, ...) {
	char msg[448];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(msg, sizeof(msg), fmt, ap);
	va_end(ap);
	snprintf(g_err, sizeof(g_err), "%s() error: %s", func, msg);
	fprintf(stderr, "%s\n", g_err);
}

int sf_hip_check(hipError_t e, const char* func, const char* what) {
	if (e == hipSuccess) return 0;
	sf_error(func, "%s failed: %s", what, hipGetErrorString(e));
	return -1;
}
