/*
 * prof.c -- per-kernel timing with HIP events recorded on the stream each
 * launch goes to (the analogue of the reference's TIMING_START/STOP macros,
 * src/timing.h:17-37, which time host phases with CLOCK_MONOTONIC).
 *
 * Disabled by default.  When enabled, each launcher brackets its kernel with
 * an event pair; sf_prof_query folds finished pairs into per-name totals.
 */
#include <stdio.h>
#include <string.h>

#include "../sf_internal.h"

#define SF_PROF_PAIRS 2048
#define SF_PROF_NAMES 64

typedef struct {
	hipEvent_t e0, e1;
	int name;
} sf_prof_rec;

static int g_on = 0;
static sf_prof_rec g_rec[SF_PROF_PAIRS];
static int g_nrec = 0;
static int g_created = 0;
static const char* g_names[SF_PROF_NAMES];
static double g_ms[SF_PROF_NAMES];
static unsigned long g_cnt[SF_PROF_NAMES];
static int g_nnames = 0;

static int name_index(const char* name) {
	for (int k = 0; k < g_nnames; ++k)
		if (strcmp(g_names[k], name) == 0) return k;
	if (g_nnames == SF_PROF_NAMES) return -1;
	g_names[g_nnames] = name;
	g_ms[g_nnames] = 0.0;
	g_cnt[g_nnames] = 0;
	return g_nnames++;
}

static void flush(void) {
	for (int k = 0; k < g_nrec; ++k) {
		float ms = 0.0f;
		if (hipEventSynchronize(g_rec[k].e1) == hipSuccess &&
		    hipEventElapsedTime(&ms, g_rec[k].e0, g_rec[k].e1) == hipSuccess && g_rec[k].name >= 0) {
			g_ms[g_rec[k].name] += (double)ms;
			g_cnt[g_rec[k].name] += 1;
		}
	}
	g_nrec = 0;
}

void sf_prof_enable(int on) {
	if (!on && g_on) flush();
	g_on = on;
}

void sf_prof_reset(void) {
	flush();
	for (int k = 0; k < g_nnames; ++k) {
		g_ms[k] = 0.0;
		g_cnt[k] = 0;
	}
}

int sf_prof_begin(const char* name, hipStream_t s) {
	if (!g_on) return -1;
	if (g_nrec == SF_PROF_PAIRS) flush();
	if (g_nrec >= g_created) {
		if (hipEventCreate(&g_rec[g_nrec].e0) != hipSuccess) return -1;
		if (hipEventCreate(&g_rec[g_nrec].e1) != hipSuccess) {
			hipEventDestroy(g_rec[g_nrec].e0);
			return -1;
		}
		g_created = g_nrec + 1;
	}
	const int tok = g_nrec;
	g_rec[tok].name = name_index(name);
	if (hipEventRecord(g_rec[tok].e0, s) != hipSuccess) return -1;
	/* taken now, so brackets may nest; an unfinished pair is flushed as if
	   it ended at its end event's last recording */
	g_nrec = tok + 1;
	if (hipEventRecord(g_rec[tok].e1, s) != hipSuccess) return -1;
	return tok;
}

void sf_prof_end(int token, hipStream_t s) {
	if (token < 0 || token >= g_nrec) return;
	hipEventRecord(g_rec[token].e1, s);
}

int sf_prof_query(const char* name, double* total_ms, unsigned long* count) {
	flush();
	for (int k = 0; k < g_nnames; ++k) {
		if (strcmp(g_names[k], name) == 0) {
			if (total_ms) *total_ms = g_ms[k];
			if (count) *count = g_cnt[k];
			return STARFLOOD_SUCCESS;
		}
	}
	if (total_ms) *total_ms = 0.0;
	if (count) *count = 0;
	return STARFLOOD_SUCCESS;
}
