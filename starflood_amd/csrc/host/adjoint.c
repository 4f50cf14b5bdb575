/*
 * adjoint.c -- forward KDK with checkpoints and its discrete adjoint, and
 * the force / force-VJP entry points on device arrays (no reference
 * counterpart).
 *
 * Forward (the fast policy's step, src/simulation.c:381-424 in structure):
 *     v += h a(x);  x64 += dt v, x = (float) x64;  a = a(x);  v += h a
 * with h = dt/2, FP32 kicks and an FP64 drift.  The tape keeps (x64, v, a)
 * every ckpt_every steps, and the positions of the last segment.
 *
 * Backward, for cotangents (lx, lv) of (x_K, v_K): the kick that ends step
 * n-1 and the kick that starts step n both act at x_n with the same lv, so
 * their VJPs merge into one with coefficient dt (h at x_0 and x_K):
 *     lx += J(x_K)^T (h lv)
 *     for n = K-1 .. 0:  lv += dt lx;  lx += J(x_n)^T (c_n lv), c_n = dt, c_0 = h
 * One force VJP per step; segments are recomputed from their checkpoints in
 * reverse order (the last one is still on the tape), so the recomputed
 * positions are bit-identical to the forward's.
 */
#include <stdlib.h>
#include <string.h>

#include "sf_host.h"

/* kernels in device/adjoint.hip */
int sfd_force_vjp(const float* body, const float* mas, const float* mu, unsigned int N, unsigned int n_pad,
                  float eps2, double G, int acc_phys, float coef, float* nu, float* part, unsigned int slices,
                  int accumulate, float* g, hipStream_t s);
int sfd_axpy(float* y, const float* x, unsigned int n, float a, hipStream_t s);
int sfd_f32_to_f64(const float* x, unsigned int n, double* y, hipStream_t s);
int sfd_f64_to_f32(const double* x, unsigned int n, float* y, hipStream_t s);
int sfd_fill0(float* y, unsigned int n, hipStream_t s);

/* device scratch for one N: packed bodies, VJP weights, slice partials */
typedef struct {
	unsigned int N, n_pad, slices;
	float* body;   /* n_pad float4 */
	float* nu;     /* n_pad float4 */
	float* part;   /* slices x 2 x n_pad float4 */
} sf_pp_scratch;

static void scratch_free(sf_pp_scratch* w) {
	if (w->body) hipFree(w->body);
	if (w->nu) hipFree(w->nu);
	if (w->part) hipFree(w->part);
	memset(w, 0, sizeof(*w));
}

static int scratch_alloc(sf_pp_scratch* w, unsigned int N, const char* func) {
	w->N = N;
	w->n_pad = ((N ? N : 1u) + 1023u) / 1024u * 1024u;
	w->slices = sf_pp_slice_count(w->n_pad);
	const size_t plane = (size_t)w->n_pad * 4u * sizeof(float);
	if (hipMalloc((void**)&w->body, SF_BODY_BYTES(w->n_pad)) != hipSuccess ||
	    hipMalloc((void**)&w->nu, plane) != hipSuccess ||
	    hipMalloc((void**)&w->part, plane * 2u * w->slices) != hipSuccess) {
		sf_error(func, "device allocation for N = %u failed", N);
		scratch_free(w);
		return -1;
	}
	/* the packed-body state words start at zero (set-up only, then synchronised) */
	if (hipMemset(w->body, 0, SF_BODY_BYTES(w->n_pad)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
		sf_error(func, "clearing the body buffer for N = %u failed", N);
		scratch_free(w);
		return -1;
	}
	return 0;
}

static float eps2_of(const sf_params_t* p) { return (float)(p->epsilon * p->epsilon); }

static int force(const sf_params_t* p, sf_pp_scratch* w, const float* x, const float* m, float* acc, float* pot,
                 hipStream_t s) {
	return sf_pp_force_arrays(p, w->N, w->n_pad, x, m, w->body, w->part,
	                          (size_t)w->slices * 2u * w->n_pad * 4u * sizeof(float), acc, pot, s);
}

/* g (=, or += when accumulate) J(x)^T (coef mu) */
static int vjp(const sf_params_t* p, sf_pp_scratch* w, const float* x, const float* m, const float* mu, float coef,
               int accumulate, float* g, hipStream_t s) {
	if (sfd_pack_bodies(x, m, w->N, w->n_pad, w->body, s)) return -1;
	return sfd_force_vjp(w->body, m, mu, w->N, w->n_pad, eps2_of(p), p->G, p->acc_mode == SF_ACC_PHYSICAL, coef, w->nu,
	                     w->part, w->slices, accumulate, g, s);
}

struct sf_tape {
	sf_params_t p;              /* epsilon, G, dt, acc_mode at the forward call */
	unsigned int N, nsteps, ckpt, n_ckpt, last_start;
	sf_pp_scratch w;
	float* m;                   /* N (copy) */
	double* ck_x64;             /* n_ckpt x 3N */
	float* ck_v;                /* n_ckpt x 3N */
	float* ck_a;                /* n_ckpt x 3N */
	float* seg;                 /* (ckpt + 1) x 3N positions of one segment */
	unsigned int seg_first;     /* step of seg[0] */
	float *x, *v, *a, *lx, *lv; /* 3N each */
	double* x64;                /* 3N */
	hipStream_t stream;         /* the stream of the tape's last call */
	hipEvent_t done;            /* recorded there by sf_tape_free */
	int have_done;
};

/* Device memory of a tape, released (hipFree waits for the device). */
static void tape_release(sf_tape* t) {
	void* ptrs[] = {t->m, t->ck_x64, t->ck_v, t->ck_a, t->seg, t->x, t->v, t->a, t->lx, t->lv, t->x64};
	for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k)
		if (ptrs[k]) hipFree(ptrs[k]);
	scratch_free(&t->w);
	if (t->have_done) hipEventDestroy(t->done);
	free(t);
}

/* The last freed tape, kept with its device memory for the next forward of
   the same N, step count and checkpoint interval (a training loop's every
   iteration after the first): no allocation, and no host synchronisation,
   only a device-side wait on the event recorded when it was freed. */
static sf_tape* g_spare;

void sf_tape_free(sf_tape* t) {
	if (!t) return;
	if (g_spare) tape_release(g_spare);
	g_spare = NULL;
	if (t->have_done && hipEventRecord(t->done, t->stream) == hipSuccess) {
		g_spare = t;
		return;
	}
	tape_release(t);
}

/* the caller's stream; NULL is HIP's default (null) stream, which is also
   PyTorch's default stream, so work stays ordered with the caller's */
static hipStream_t stream_of(void* stream) { return (hipStream_t)stream; }

/* one KDK step from (x64, x, v, a) in place; a is the force at x on entry */
static int kdk(sf_tape* t, hipStream_t s) {
	const unsigned int n3 = 3u * t->N;
	const float h = 0.5f * (float)t->p.dt;
	if (sfd_kick_drift_fast(t->v, t->a, t->x64, t->x, n3, h, t->p.dt, 0, s)) return -1;
	if (force(&t->p, &t->w, t->x, t->m, t->a, NULL, s)) return -1;
	return sfd_kick_fast(t->v, t->a, n3, h, s);
}

static int forward_run(sf_tape* t, const float* x0, const float* v0, const float* m, float* xN, float* vN,
                       hipStream_t s, const char* fn);

int sf_adjoint_forward(unsigned int N, const float* x0, const float* v0, const float* m, unsigned int nsteps,
                       unsigned int ckpt_every, float* xN, float* vN, void* stream, sf_tape** tape_out) {
	const char* fn = "sf_adjoint_forward";
	if (tape_out) *tape_out = NULL;
	if (!x0 || !v0 || !m || !xN || !vN || !tape_out || N == 0 || ckpt_every == 0) {
		sf_error(fn, "invalid arguments (null array, N = 0 or ckpt_every = 0)");
		return STARFLOOD_FAILURE;
	}
	hipStream_t s = stream_of(stream);
	const unsigned int ckpt = ckpt_every < nsteps ? ckpt_every : (nsteps ? nsteps : 1u);
	const size_t n3 = 3u * (size_t)N, f3 = n3 * sizeof(float);
	sf_tape* t = g_spare;
	if (t && t->N == N && t->nsteps == nsteps && t->ckpt == ckpt) {
		/* the spare tape's buffers, once its last call's work is done */
		g_spare = NULL;
		int rc = hipStreamWaitEvent(s, t->done, 0) != hipSuccess;
		t->p = *sf_params();
		t->stream = s;
		if (!rc) rc = forward_run(t, x0, v0, m, xN, vN, s, fn);
		if (rc) {
			tape_release(t);
			return STARFLOOD_FAILURE;
		}
		*tape_out = t;
		return STARFLOOD_SUCCESS;
	}
	t = (sf_tape*)calloc(1, sizeof(sf_tape));
	if (!t) return STARFLOOD_FAILURE;
	t->p = *sf_params();
	t->stream = s;
	t->N = N;
	t->nsteps = nsteps;
	t->ckpt = ckpt;
	t->n_ckpt = nsteps ? (nsteps + t->ckpt - 1u) / t->ckpt : 0u;
	t->last_start = t->n_ckpt ? (t->n_ckpt - 1u) * t->ckpt : 0u;
	t->have_done = hipEventCreateWithFlags(&t->done, hipEventDisableTiming) == hipSuccess;
	int rc = scratch_alloc(&t->w, N, fn);
	rc |= hipMalloc((void**)&t->m, (size_t)N * sizeof(float)) != hipSuccess;
	rc |= hipMalloc((void**)&t->ck_x64, (t->n_ckpt ? t->n_ckpt : 1u) * n3 * sizeof(double)) != hipSuccess;
	rc |= hipMalloc((void**)&t->ck_v, (t->n_ckpt ? t->n_ckpt : 1u) * f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->ck_a, (t->n_ckpt ? t->n_ckpt : 1u) * f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->seg, (t->ckpt + 1u) * f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->x, f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->v, f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->a, f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->lx, f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->lv, f3) != hipSuccess;
	rc |= hipMalloc((void**)&t->x64, n3 * sizeof(double)) != hipSuccess;
	if (rc) {
		sf_error(fn, "device allocation for N = %u, %u steps failed", N, nsteps);
		tape_release(t);
		return STARFLOOD_FAILURE;
	}
	if (forward_run(t, x0, v0, m, xN, vN, s, fn)) {
		tape_release(t);
		return STARFLOOD_FAILURE;
	}
	*tape_out = t;
	return STARFLOOD_SUCCESS;
}

/* The forward integration into tape t (buffers in place) on stream s. */
static int forward_run(sf_tape* t, const float* x0, const float* v0, const float* m, float* xN, float* vN,
                       hipStream_t s, const char* fn) {
	const unsigned int N = t->N, nsteps = t->nsteps;
	const size_t n3 = 3u * (size_t)N, f3 = n3 * sizeof(float);
	int rc = 0;
	rc |= hipMemcpyAsync(t->m, m, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(t->x, x0, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(t->v, v0, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= sfd_f32_to_f64(t->x, (unsigned int)n3, t->x64, s);
	const int tok = sf_prof_begin("adjoint_forward", s);
	rc |= force(&t->p, &t->w, t->x, t->m, t->a, NULL, s);
	for (unsigned int n = 0; n < nsteps && !rc; ++n) {
		if (n % t->ckpt == 0u) {
			const size_t c = n / t->ckpt;
			rc |= hipMemcpyAsync(t->ck_x64 + c * n3, t->x64, n3 * sizeof(double), hipMemcpyDeviceToDevice, s) !=
			      hipSuccess;
			rc |= hipMemcpyAsync(t->ck_v + c * n3, t->v, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
			rc |= hipMemcpyAsync(t->ck_a + c * n3, t->a, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
		}
		if (n >= t->last_start)
			rc |= hipMemcpyAsync(t->seg + (n - t->last_start) * n3, t->x, f3, hipMemcpyDeviceToDevice, s) !=
			      hipSuccess;
		rc |= kdk(t, s);
	}
	if (nsteps && !rc)
		rc |= hipMemcpyAsync(t->seg + (nsteps - t->last_start) * n3, t->x, f3, hipMemcpyDeviceToDevice, s) !=
		      hipSuccess;
	sf_prof_end(tok, s);
	t->seg_first = t->last_start;
	rc |= hipMemcpyAsync(xN, t->x, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(vN, t->v, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	if (rc || sf_hip_check(hipGetLastError(), fn, "launch")) {
		sf_error(fn, "forward integration failed");
		return -1;
	}
	return 0;
}

/* The positions of segment c that the backward reads, into t->seg:
   x_{start} .. x_{start+len-1}, and x_{start+len} too for the last segment.
   Only the last segment reads its end position, so the others take len - 1
   steps (none with a checkpoint every step).  The forward leaves the last
   segment on the tape; it is recomputed only when the tape's backward runs
   again. */
static int recompute(sf_tape* t, unsigned int c, unsigned int len, hipStream_t s) {
	const size_t n3 = 3u * (size_t)t->N, f3 = n3 * sizeof(float);
	const unsigned int steps = c + 1u == t->n_ckpt ? len : len - 1u;
	int rc = 0;
	t->seg_first = c * t->ckpt;
	if (steps == 0u) return sfd_f64_to_f32(t->ck_x64 + c * n3, (unsigned int)n3, t->seg, s);
	rc |= hipMemcpyAsync(t->x64, t->ck_x64 + c * n3, n3 * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(t->v, t->ck_v + c * n3, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(t->a, t->ck_a + c * n3, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= sfd_f64_to_f32(t->x64, (unsigned int)n3, t->x, s);
	rc |= hipMemcpyAsync(t->seg, t->x, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	for (unsigned int k = 1; k <= steps && !rc; ++k) {
		rc |= kdk(t, s);
		rc |= hipMemcpyAsync(t->seg + k * n3, t->x, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	}
	return rc;
}

int sf_adjoint_backward(sf_tape* t, const float* gxN, const float* gvN, float* gx0, float* gv0, void* stream) {
	const char* fn = "sf_adjoint_backward";
	if (!t || !gx0 || !gv0) {
		sf_error(fn, "invalid arguments (null tape or output)");
		return STARFLOOD_FAILURE;
	}
	hipStream_t s = stream_of(stream);
	t->stream = s;
	const unsigned int n3 = 3u * t->N;
	const size_t f3 = (size_t)n3 * sizeof(float);
	const float dt = (float)t->p.dt, h = 0.5f * dt;
	int rc = 0;
	rc |= gxN ? hipMemcpyAsync(t->lx, gxN, f3, hipMemcpyDeviceToDevice, s) != hipSuccess : sfd_fill0(t->lx, n3, s);
	rc |= gvN ? hipMemcpyAsync(t->lv, gvN, f3, hipMemcpyDeviceToDevice, s) != hipSuccess : sfd_fill0(t->lv, n3, s);
	const int tok = sf_prof_begin("adjoint_backward", s);
	for (unsigned int c = t->n_ckpt; c-- > 0u && !rc;) {
		const unsigned int start = c * t->ckpt;
		const unsigned int len = (t->nsteps - start) < t->ckpt ? (t->nsteps - start) : t->ckpt;
		if (t->seg_first != start) rc |= recompute(t, c, len, s);
		/* the kick that ends the last step acts at x_K alone */
		if (c == t->n_ckpt - 1u) rc |= vjp(&t->p, &t->w, t->seg + (size_t)len * n3, t->m, t->lv, h, 1, t->lx, s);
		for (unsigned int k = len; k-- > 0u && !rc;) {
			const unsigned int n = start + k;
			rc |= sfd_axpy(t->lv, t->lx, n3, dt, s);
			rc |= vjp(&t->p, &t->w, t->seg + (size_t)k * n3, t->m, t->lv, n > 0u ? dt : h, 1, t->lx, s);
		}
	}
	sf_prof_end(tok, s);
	rc |= hipMemcpyAsync(gx0, t->lx, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	rc |= hipMemcpyAsync(gv0, t->lv, f3, hipMemcpyDeviceToDevice, s) != hipSuccess;
	if (rc || sf_hip_check(hipGetLastError(), fn, "launch")) {
		sf_error(fn, "adjoint integration failed");
		return STARFLOOD_FAILURE;
	}
	return STARFLOOD_SUCCESS;
}

/* single evaluations on caller arrays; the scratch is kept between calls
   (one per process: these two entry points are not reentrant) */
static sf_pp_scratch g_w;

static int scratch_for(unsigned int N, const char* fn) {
	if (g_w.body && g_w.N == N) return 0;
	if (g_w.body) {
		hipDeviceSynchronize();
		scratch_free(&g_w);
	}
	return scratch_alloc(&g_w, N, fn);
}

int sf_force_dev(unsigned int N, const float* x, const float* m, float* acc, float* pot, void* stream) {
	const char* fn = "sf_force_dev";
	if (N == 0) return STARFLOOD_SUCCESS;
	if (!x || !m || !acc) {
		sf_error(fn, "null array");
		return STARFLOOD_FAILURE;
	}
	if (scratch_for(N, fn)) return STARFLOOD_FAILURE;
	hipStream_t s = stream_of(stream);
	if (force(sf_params(), &g_w, x, m, acc, pot, s)) return STARFLOOD_FAILURE;
	return sf_hip_check(hipStreamSynchronize(s), fn, "sync") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

int sf_force_vjp_dev(unsigned int N, const float* x, const float* m, const float* mu, float* g, void* stream) {
	const char* fn = "sf_force_vjp_dev";
	if (N == 0) return STARFLOOD_SUCCESS;
	if (!x || !m || !mu || !g) {
		sf_error(fn, "null array");
		return STARFLOOD_FAILURE;
	}
	if (scratch_for(N, fn)) return STARFLOOD_FAILURE;
	hipStream_t s = stream_of(stream);
	if (vjp(sf_params(), &g_w, x, m, mu, 1.0f, 0, g, s)) return STARFLOOD_FAILURE;
	return sf_hip_check(hipStreamSynchronize(s), fn, "sync") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}
