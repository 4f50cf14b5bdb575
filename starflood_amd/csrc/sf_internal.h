/*
 * sf_internal.h -- interface between the C host layer (csrc/host/ C files,
 * compiled by gcc) and the HIP device layer (csrc/device/ HIP files, compiled by
 * hipcc for gfx950).  Not installed; callers use include/starflood*.h.
 *
 * Device launchers take device pointers and a stream, enqueue work and
 * return 0 or -1; they never synchronise unless their comment says so.
 */
#ifndef SF_INTERNAL_H
#define SF_INTERNAL_H

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__
#endif
#include <hip/hip_runtime_api.h>

#include "../../include/starflood_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------- error reporting (host/params.c) ---------- */
/* Record and print "<func>() error: <msg>" in the reference's style. */
void sf_error(const char* func, const char* fmt, ...);
/* Check a HIP call; on failure record the error and return -1. */
int sf_hip_check(hipError_t e, const char* func, const char* what);
#define SF_HIP(call, func) do { if (sf_hip_check((call), (func), #call)) return STARFLOOD_FAILURE; } while (0)

/* Current process-wide parameters (host/params.c). */
const sf_params_t* sf_params(void);

/* ---------- per-simulation device context (host/context.c) ---------- */
typedef struct sf_tree_dev sf_tree_dev;   /* host/tree.c */

typedef struct sf_ctx {
	void* host_mem;          /* key: sim_t.mem */
	void* host_pinned_base;  /* the page-locked allocation holding sim_t.mem, or NULL (ordinary memory) */
	unsigned int N;
	unsigned int dof;        /* reals per particle in sim.mem */
	unsigned int n_pad;      /* N rounded up to the PP j tile */
	float* d_mem;            /* device copy of sim.mem, same layout */
	float* d_body;           /* SF_BODY_BYTES(n_pad): x, y, z, m (pad: m = 0), then SF_BODY_STATE words */
	double* d_x64;           /* 3N FP64 positions for the FAST drift */
	float* d_part;           /* PP partial sums, 8 * slices * n_pad floats */
	unsigned int part_slices;
	double* d_red;           /* reduction scratch (blocks * SF_RED_VALUES) */
	double* h_red;           /* pinned host mirror of the reduction result */
	float* d_u;              /* SPH specific internal energy (N) */
	float* d_dudt;           /* its derivative (N) */
	float* d_sph_tmp;        /* SPH scratch (packed records, 12 n_pad floats) */
	void* d_sph_grid;        /* cell-grid pair search scratch (cubic model) */
	hipStream_t stream;
	sf_stats_t stats;        /* statistics of the last sim_step */
	int have_u;
} sf_ctx;

sf_ctx* sf_ctx_lookup(const void* host_mem, const char* func);
sf_ctx* sf_ctx_create(void* host_mem, unsigned int N, unsigned int dof);
void sf_ctx_destroy(sf_ctx* c);
/* Ensure the PP partial-sum scratch holds `slices` slices. */
int sf_ctx_reserve_part(sf_ctx* c, unsigned int slices);
/* allocates (zeroed) d_u and d_dudt on first use */
int sf_ctx_ensure_u(sf_ctx* c);
/* The library's stream for work not tied to a simulation (lazy). */
hipStream_t sf_default_stream(void);

/* block offsets into d_mem / host mem (reference src/simulation.c:33-53) */
static inline float* sf_blk(float* base, unsigned int N, unsigned int e) { return base + (size_t)N * (size_t)e; }

#define SF_E_POS 0u
#define SF_E_VEL 3u
#define SF_E_ACC 6u
#define SF_E_MAS 9u
#define SF_E_POT 10u
#define SF_E_KEN 11u
#define SF_E_PEN 12u
#define SF_E_RAD 13u
#define SF_E_RHO 14u
#define SF_E_PRS 15u

/* mirror-mode transfers of blocks [e0, e0+count) (no-ops in RESIDENT mode) */
int sf_upload_blocks(sf_ctx* c, unsigned int e0, unsigned int count);
int sf_download_blocks(sf_ctx* c, unsigned int e0, unsigned int count);
int sf_finish(sf_ctx* c, const char* func);   /* synchronise the stream, report errors */

/* ---------- kernel timing (host/prof.c) ---------- */
/* Bracket a launch: sf_prof_begin returns a token passed to sf_prof_end. */
int sf_prof_begin(const char* name, hipStream_t s);
void sf_prof_end(int token, hipStream_t s);

/* ---------- device launchers (csrc/device) ---------- */
#define SF_PP_TILE 256u      /* j padding granule of d_body */
/* packed bodies: n_pad float4, then SF_BODY_STATE words that sfd_pack_bodies
   reads and writes (the leading bodies that form whole tiles of equal mass,
   and two block-count words that must be zero when the buffer is made) */
#define SF_BODY_STATE 64u
#define SF_BODY_BYTES(n_pad) ((size_t)(n_pad) * 16u + SF_BODY_STATE * 4u)
#define SF_RED_VALUES 16u    /* doubles per reduction partial */

/* body[i] = (pos_i, m_i) for i < N, zero for N <= i < n_pad */
int sfd_pack_bodies(const float* pos, const float* mas, unsigned int N, unsigned int n_pad,
                    float* body, hipStream_t s);

/* Reference-order PP gravity (src/gravity.c:693-812), one lane per i. */
int sfd_pp_exact(const float* pos, const float* mas, unsigned int N, float eps2, float negG,
                 int acc_phys, int kahan_acc, int kahan_pot, float* acc, float* pot, hipStream_t s);

/* Fast PP gravity: j-sliced partial sums into part (sfd_pp_part_bytes of
   n_pad and slices), then a fixed-order FP64 combine into acc (AoS) and pot.
   pot may be NULL (skip). */
int sfd_pp_fast(const float* body, unsigned int N, unsigned int n_pad, float eps2, double G,
                int acc_phys, float* part, unsigned int slices, float* acc, float* pot, hipStream_t s);
size_t sfd_pp_part_bytes(unsigned int n_pad, unsigned int slices);

/* FP64 direct sum for the listed particles (device idx list). */
int sfd_pp_f64(const float* pos, const float* mas, unsigned int N, double eps2, double G, int acc_phys,
               const unsigned int* idx, unsigned int n_idx, double* acc, double* pot, hipStream_t s);

/* Reference KDK pieces (src/simulation.c:381-451), FP32, no contraction. */
int sfd_kick_exact(float* vel, const float* acc, unsigned int n3, float half_dt, hipStream_t s);
int sfd_drift_exact(float* pos, const float* vel, unsigned int n3, float dt, hipStream_t s);
/* ken_i = (0.5 m_i) |v_i|^2, pen_i = m_i pot_i */
int sfd_energy_terms(const float* vel, const float* mas, const float* pot, unsigned int N,
                     float* ken, float* pen, hipStream_t s);
/* Sequential FP64 Kahan of x[0..N) in index order (src/simulation.c:455-497),
   written to *out.  One lane; for bit parity. */
int sfd_energy_sum_sequential(const float* x, unsigned int N, double* out, hipStream_t s);

/* FAST KDK: kick + FP64 drift fused; x64 refreshed from pos where they
   disagree (sfd_x64_refresh, or inside the drift with refresh = 1). */
int sfd_x64_refresh(const float* pos, double* x64, unsigned int n3, hipStream_t s);
int sfd_kick_drift_fast(float* vel, const float* acc, double* x64, float* pos, unsigned int n3,
                        float half_dt, double dt, int refresh, hipStream_t s);
int sfd_kick_fast(float* vel, const float* acc, unsigned int n3, float half_dt, hipStream_t s);
/* sfd_kick_fast then sfd_energy_terms in one pass over the particles */
int sfd_kick_energy_fast(float* vel, const float* acc, const float* mas, const float* pot, unsigned int N,
                         float half_dt, float* ken, float* pen, hipStream_t s);
/* sph.hip: u += half_dt du/dt (cubic SPH model) */
int sfd_sph_u_kick(float* u, const float* dudt, unsigned int N, float half_dt, hipStream_t s);

/* Deterministic FP64 reductions over the state: fills out[0..SF_RED_VALUES)
   with T, V, W, U, Px, Py, Pz, Lx, Ly, Lz.  x64 may be NULL (use pos).
   u may be NULL. red is scratch of sfd_reduce_scratch_doubles(N) doubles. */
unsigned int sfd_reduce_scratch_doubles(unsigned int N);
int sfd_reduce_state(const float* pos, const double* x64, const float* vel, const float* mas,
                     const float* pot, const float* ken, const float* pen, const float* u,
                     unsigned int N, double G, double eps, double* red, hipStream_t s);

#ifdef __cplusplus
}
#endif

#endif /* SF_INTERNAL_H */
