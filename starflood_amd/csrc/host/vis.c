/*
 * vis.c -- the reference's visualisation entry points (src/visualization.h)
 * over the device rasteriser of device/vis.hip.
 *
 * vis_t and its host buffers are laid out as in the reference.  The device
 * keeps the accumulation and render buffers of each vis_t (made by its first
 * vis_draw), which copies both back into the host buffers.  The camera
 * matrix is computed on the host with the reference's float operations and
 * the host libm (src/visualization.c:330-486); the mean radius it scales by
 * (AUTOSCALING) is the reference's sequential FP64 Kahan sum, on the device.
 */
#define _POSIX_C_SOURCE 200809L

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/starflood_vis.h"
#include "../device/tree_dev.h"
#include "sf_host.h"

/* device/vis.hip */
int sfd_vis_draw(const float* pos, const float* vel, const float* pot, unsigned int N, const float mvp[16],
                 unsigned int step, float dt, unsigned int sizex, unsigned int sizey, double scale, int* buf,
                 float* render, hipStream_t s);

#define SF_VIS_ALIGN 4194304u  /* STARFLOOD_ALIGNMENT, config.h:230 */
#define SF_VIS_SAMPLES 128      /* SPATIAL_SAMPLES, config.h:203 */
#define SF_VIS_EXPOSURE (-1.000) /* EXPOSURE, config.h:206 */
#define SF_VIS_ORTHO (-1.000)    /* ORTHO_SCALE, config.h:209 */
#define SF_VIS_MAX 8
#define SF_TAU 6.28318530717958647692528676655900577

typedef struct {
	void* key;            /* vis_t.mem */
	unsigned int sizex, sizey;
	int* d_atomic;
	float* d_render;
	double* d_r1;         /* radii for the mean radius */
	unsigned int r1_n;
	double* d_avg;
	double* h_avg;        /* pinned */
} sf_vis_dev;

static sf_vis_dev* g_vis[SF_VIS_MAX];

/* vis_init's page-locked buffers (the device copies back into them at full
   PCIe rate): the aligned address handed out, and the allocation to free */
typedef struct {
	void* aligned;
	void* base;
} sf_vis_pinned;
static sf_vis_pinned g_pinned[SF_VIS_MAX];

static void vis_dev_destroy(sf_vis_dev* d) {
	if (!d) return;
	for (int k = 0; k < SF_VIS_MAX; ++k)
		if (g_vis[k] == d) g_vis[k] = NULL;
	if (d->d_atomic) hipFree(d->d_atomic);
	if (d->d_render) hipFree(d->d_render);
	if (d->d_r1) hipFree(d->d_r1);
	if (d->d_avg) hipFree(d->d_avg);
	if (d->h_avg) hipHostFree(d->h_avg);
	free(d);
}

static sf_vis_dev* vis_dev_get(const vis_t* v) {
	for (int k = 0; k < SF_VIS_MAX; ++k)
		if (g_vis[k] && g_vis[k]->key == v->mem) return g_vis[k];
	int slot = -1;
	for (int k = 0; k < SF_VIS_MAX; ++k)
		if (!g_vis[k]) {
			slot = k;
			break;
		}
	sf_vis_dev* d = slot >= 0 ? (sf_vis_dev*)calloc(1, sizeof(sf_vis_dev)) : NULL;
	if (!d) {
		sf_error("vis_draw", "too many live visualisations");
		return NULL;
	}
	d->key = v->mem;
	d->sizex = v->sizex;
	d->sizey = v->sizey;
	const size_t n = 4u * (size_t)v->sizex * v->sizey;
	if (hipMalloc((void**)&d->d_atomic, (n ? n : 1u) * sizeof(int)) != hipSuccess ||
	    hipMalloc((void**)&d->d_render, (n ? n : 1u) * sizeof(float)) != hipSuccess ||
	    hipMalloc((void**)&d->d_avg, sizeof(double)) != hipSuccess ||
	    hipHostMalloc((void**)&d->h_avg, sizeof(double), hipHostMallocDefault) != hipSuccess) {
		sf_error("vis_draw", "device allocation for %u x %u pixels failed", v->sizex, v->sizey);
		vis_dev_destroy(d);
		return NULL;
	}
	g_vis[slot] = d;
	return d;
}

/* reference src/visualization.c:37-190 */
int vis_init(vis_t* SF_RESTRICT vis_ptr, const unsigned int sizex, const unsigned int sizey) {
	if (!vis_ptr) return STARFLOOD_FAILURE;
	const size_t px = (size_t)sizex * sizey;
	/* atomic (i32 RGBA), render (f32 RGBA), binary (room for f32 RGB) */
	const size_t mem_size = 16u * px + 16u * px + 12u * px;
	void* mem = NULL;
	/* page-locked when a device is present and a slot is free, else
	   ordinary host memory (as the reference's aligned_alloc) */
	int slot = -1;
	for (int k = 0; k < SF_VIS_MAX && slot < 0; ++k)
		if (!g_pinned[k].base) slot = k;
	void* base = NULL;
	if (slot >= 0 && hipHostMalloc(&base, mem_size + SF_VIS_ALIGN, hipHostMallocDefault) == hipSuccess && base) {
		mem = (void*)(((uintptr_t)base + (SF_VIS_ALIGN - 1u)) & ~(uintptr_t)(SF_VIS_ALIGN - 1u));
		g_pinned[slot].aligned = mem;
		g_pinned[slot].base = base;
	} else {
		(void)hipGetLastError();
		if (posix_memalign(&mem, (size_t)SF_VIS_ALIGN, mem_size ? mem_size : 1u) != 0 || !mem) {
			fprintf(stderr, "%s error: mem is NULL after posix_memalign(&mem, %zu, %zu)\n", "vis_init()",
			        (size_t)SF_VIS_ALIGN, mem_size);
			return STARFLOOD_FAILURE;
		}
	}
	memset(mem, 0, mem_size ? mem_size : 1u);
	vis_ptr->sizex = sizex;
	vis_ptr->sizey = sizey;
	vis_ptr->mem = mem;
	vis_ptr->atomic_buffer = (i32*)mem;
	vis_ptr->render_buffer = (f32*)((char*)mem + 16u * px);
	vis_ptr->binary_buffer = (void*)((char*)mem + 32u * px);
	return STARFLOOD_SUCCESS;
}

/* reference src/visualization.c:192-230 */
int vis_free(vis_t* SF_RESTRICT vis_ptr) {
	if (!vis_ptr) return STARFLOOD_FAILURE;
	if (vis_ptr->mem) {
		for (int k = 0; k < SF_VIS_MAX; ++k)
			if (g_vis[k] && g_vis[k]->key == vis_ptr->mem) {
				hipDeviceSynchronize();
				vis_dev_destroy(g_vis[k]);
			}
	}
	int pinned = 0;
	for (int k = 0; k < SF_VIS_MAX && vis_ptr->mem; ++k)
		if (g_pinned[k].aligned == vis_ptr->mem) {
			hipHostFree(g_pinned[k].base);
			g_pinned[k].aligned = g_pinned[k].base = NULL;
			pinned = 1;
		}
	if (!pinned) free(vis_ptr->mem);
	vis_ptr->sizex = 0u;
	vis_ptr->sizey = 0u;
	vis_ptr->mem = NULL;
	vis_ptr->atomic_buffer = NULL;
	vis_ptr->render_buffer = NULL;
	vis_ptr->binary_buffer = NULL;
	return STARFLOOD_SUCCESS;
}

/* C = A * B in the reference's term order (src/visualization.c:456-486) */
static void mat_mul(const float* A, const float* B, float* out) {
	float C[16];
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			C[4 * i + j] = (A[4 * i + 3] * B[4 * 3 + j]) + (A[4 * i + 2] * B[4 * 2 + j]) +
			               (A[4 * i + 1] * B[4 * 1 + j]) + (A[4 * i + 0] * B[4 * 0 + j]);
	memcpy(out, C, sizeof(C));
}

/* model-view-projection of src/visualization.c:330-486 at time t, with the
   AUTOSCALING mean radius */
static void vis_mvp(float time, double average_radius, float mvp[16]) {
	float M[16] = {0}, V[16] = {0}, P[16] = {0};
	for (int i = 0; i < 4; i++) V[5 * i] = P[5 * i] = 1.0f;
	float alp = 0.0f, bet = 0.0f, gam = 0.0f;
	bet = (float)SF_TAU * 0.1250f;
	gam = (float)SF_TAU * 0.0625f;
	alp += (float)SF_TAU * 1.000e-4f * time;
	bet += (float)SF_TAU * 1.000e-4f * time;
	gam += (float)SF_TAU * 1.000e-4f * time;
	const float ca = cosf(alp), sa = sinf(alp);
	const float cb = cosf(bet), sb = sinf(bet);
	const float cg = cosf(gam), sg = sinf(gam);
	const float rot[16] = {ca * cb, ca * sb * sg - sa * cg, ca * sb * cg + sa * sg, 0.0f,
	                       sa * cb, sa * sb * sg + ca * cg, sa * sb * cg - ca * sg, 0.0f,
	                       -sb,     cb * sg,                cb * cg,                0.0f,
	                       0.0f,    0.0f,                   0.0f,                   1.0f};
	memcpy(M, rot, sizeof(rot));
	float ortho = (float)exp2(SF_VIS_ORTHO);
	ortho *= (float)(1.0 / average_radius);
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) P[4 * i + j] *= ortho;
	float VM[16];
	mat_mul(V, M, VM);
	mat_mul(P, VM, mvp);
}

/* reference src/visualization.c:253-736 */
int vis_draw(const vis_t* SF_RESTRICT vis_ptr, const sim_t* SF_RESTRICT sim_ptr) {
	if (!vis_ptr || !sim_ptr || !vis_ptr->mem) return STARFLOOD_FAILURE;
	sf_ctx* c = sf_ctx_lookup(sim_ptr->mem, "vis_draw");
	if (!c) return STARFLOOD_FAILURE;
	sf_vis_dev* d = vis_dev_get(vis_ptr);
	if (!d) return STARFLOOD_FAILURE;
	const sf_params_t* p = sf_params();
	const unsigned int N = c->N;
	hipStream_t s = c->stream;
	/* pos, vel (blocks 0..5) and pot */
	if (sf_upload_blocks(c, SF_E_POS, 6u) || sf_upload_blocks(c, SF_E_POT, 1u)) return STARFLOOD_FAILURE;
	const float* pos = sf_blk(c->d_mem, N, SF_E_POS);
	double average_radius = 1.0;
	if (N) {
		if (d->r1_n < N) {
			SF_HIP(hipStreamSynchronize(s), "vis_draw");
			if (d->d_r1) hipFree(d->d_r1);
			d->d_r1 = NULL;
			d->r1_n = 0;
			SF_HIP(hipMalloc((void**)&d->d_r1, (size_t)N * sizeof(double)), "vis_draw");
			d->r1_n = N;
		}
		if (sfd_mean_radius(pos, N, d->d_r1, d->d_avg, s)) return STARFLOOD_FAILURE;
		SF_HIP(hipMemcpyAsync(d->h_avg, d->d_avg, sizeof(double), hipMemcpyDeviceToHost, s), "vis_draw");
		SF_HIP(hipStreamSynchronize(s), "vis_draw");
		average_radius = *d->h_avg;
	} else {
		average_radius = 0.0 / 0.0;   /* the reference divides 0 by 0 */
	}
	if (p->verbose) printf("average_radius: % .15f\n", average_radius);
	const float time = (float)(p->dt * (double)sim_ptr->step_number);
	float mvp[16];
	vis_mvp(time, average_radius, mvp);
	const double scale = 0.01 / (double)SF_VIS_SAMPLES * exp2(SF_VIS_EXPOSURE);
	if (sfd_vis_draw(pos, sf_blk(c->d_mem, N, SF_E_VEL), sf_blk(c->d_mem, N, SF_E_POT), N, mvp,
	                 sim_ptr->step_number, (float)p->dt, vis_ptr->sizex, vis_ptr->sizey, scale, d->d_atomic,
	                 d->d_render, s))
		return STARFLOOD_FAILURE;
	const size_t n = 4u * (size_t)vis_ptr->sizex * vis_ptr->sizey;
	if (n) {
		SF_HIP(hipMemcpyAsync(vis_ptr->atomic_buffer, d->d_atomic, n * sizeof(int), hipMemcpyDeviceToHost, s),
		       "vis_draw");
		SF_HIP(hipMemcpyAsync(vis_ptr->render_buffer, d->d_render, n * sizeof(float), hipMemcpyDeviceToHost, s),
		       "vis_draw");
	}
	return sf_finish(c, "vis_draw") ? STARFLOOD_FAILURE : STARFLOOD_SUCCESS;
}

/* ---------------------------------------------------------------- files */

static uint32_t g_crc_table[256];

static void crc_init(void) {
	if (g_crc_table[1]) return;
	for (uint32_t n = 0; n < 256u; ++n) {
		uint32_t c = n;
		for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
		g_crc_table[n] = c;
	}
}

static uint32_t crc_update(uint32_t crc, const unsigned char* b, size_t n) {
	for (size_t i = 0; i < n; ++i) crc = g_crc_table[(crc ^ b[i]) & 0xFFu] ^ (crc >> 8);
	return crc;
}

static int put_be32(FILE* f, uint32_t v) {
	const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
	                            (unsigned char)v};
	return fwrite(b, 1, 4, f) == 4 ? 0 : -1;
}

/* a PNG chunk: length, type, data, CRC-32 of type and data */
static int png_chunk(FILE* f, const char type[4], const unsigned char* data, size_t n) {
	uint32_t crc = crc_update(0xFFFFFFFFu, (const unsigned char*)type, 4);
	crc = crc_update(crc, data, n);
	if (put_be32(f, (uint32_t)n) || fwrite(type, 1, 4, f) != 4 || (n && fwrite(data, 1, n, f) != n) ||
	    put_be32(f, crc ^ 0xFFFFFFFFu))
		return -1;
	return 0;
}

/* 8-bit RGB PNG (RFC 2083) whose zlib stream (RFC 1950) holds stored
   deflate blocks (RFC 1951, BTYPE 00): filter type 0 on every row. */
static int write_png(const unsigned char* rgb, size_t w, size_t h, const char* filename) {
	crc_init();
	const size_t row = 3u * w + 1u, raw_n = row * h;
	const size_t nblocks = raw_n ? (raw_n + 65534u) / 65535u : 1u;
	const size_t z_n = 2u + raw_n + 5u * nblocks + 4u;
	unsigned char* z = (unsigned char*)malloc(z_n);
	if (!z) return -1;
	size_t o = 0;
	z[o++] = 0x78;   /* deflate, 32 KiB window */
	z[o++] = 0x01;   /* no preset dictionary, check bits: 0x7801 % 31 == 0 */
	uint32_t a = 1u, b = 0u;   /* Adler-32 of the raw rows */
	size_t done = 0;
	for (size_t k = 0; k < nblocks; ++k) {
		const size_t len = raw_n - done < 65535u ? raw_n - done : 65535u;
		z[o++] = (unsigned char)(k + 1u == nblocks ? 1u : 0u);
		z[o++] = (unsigned char)(len & 0xFFu);
		z[o++] = (unsigned char)(len >> 8);
		z[o++] = (unsigned char)(~len & 0xFFu);
		z[o++] = (unsigned char)((~len >> 8) & 0xFFu);
		for (size_t i = 0; i < len; ++i) {
			const size_t q = done + i, y = q / row, x = q % row;
			const unsigned char v = x == 0 ? 0u : rgb[y * 3u * w + (x - 1u)];
			z[o++] = v;
			a = (a + v) % 65521u;
			b = (b + a) % 65521u;
		}
		done += len;
	}
	const uint32_t adler = (b << 16) | a;
	z[o++] = (unsigned char)(adler >> 24);
	z[o++] = (unsigned char)(adler >> 16);
	z[o++] = (unsigned char)(adler >> 8);
	z[o++] = (unsigned char)adler;
	FILE* f = fopen(filename, "wb");
	int rc = f ? 0 : -1;
	if (f) {
		static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
		const unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
		                                (unsigned char)w, (unsigned char)(h >> 24), (unsigned char)(h >> 16),
		                                (unsigned char)(h >> 8), (unsigned char)h, 8u, 2u, 0u, 0u, 0u};
		if (fwrite(sig, 1, 8, f) != 8 || png_chunk(f, "IHDR", ihdr, 13) || png_chunk(f, "IDAT", z, o) ||
		    png_chunk(f, "IEND", NULL, 0))
			rc = -1;
		if (fclose(f) != 0) rc = -1;
	}
	free(z);
	return rc;
}

/* reference src/visualization_io.c:27-51 */
static int write_pfm(const float* rgb, size_t w, size_t h, const char* filename) {
	FILE* f = fopen(filename, "wb");
	if (!f) return -1;
	int rc = fprintf(f, "PF\n%zu %zu\n-1.0\n", w, h) > 0 ? 0 : -1;
	if (fwrite(rgb, sizeof(float), 3u * w * h, f) != 3u * w * h) rc = -1;
	if (fclose(f) != 0) rc = -1;
	return rc;
}

/* reference src/visualization_io.c:53-77 */
static int write_ppm(const unsigned char* rgb, size_t w, size_t h, const char* filename) {
	FILE* f = fopen(filename, "wb");
	if (!f) return -1;
	int rc = fprintf(f, "P6\n%zu %zu %u\n", w, h, 255u) > 0 ? 0 : -1;
	if (fwrite(rgb, 1, 3u * w * h, f) != 3u * w * h) rc = -1;
	if (fclose(f) != 0) rc = -1;
	return rc;
}

/* reference src/visualization_io.c:166-330 */
int vis_save(const vis_t* SF_RESTRICT vis_ptr, const char* SF_RESTRICT filename, image_filetype_t filetype) {
	if (!vis_ptr || !vis_ptr->mem || !filename) return STARFLOOD_FAILURE;
	if (filetype != IMAGE_FILETYPE_PFM && filetype != IMAGE_FILETYPE_PPM && filetype != IMAGE_FILETYPE_PNG) {
		sf_error("vis_save", "unknown image file type %d", (int)filetype);
		return STARFLOOD_FAILURE;
	}
	const unsigned int w = vis_ptr->sizex, h = vis_ptr->sizey;
	const f32* render = vis_ptr->render_buffer;
	int rc;
	if (filetype == IMAGE_FILETYPE_PFM) {
		f32* bin = (f32*)vis_ptr->binary_buffer;
		for (unsigned int y = 0u; y < h; y++)
			for (unsigned int x = 0u; x < w; x++)
				for (unsigned int c = 0u; c < 3u; c++) bin[3u * (w * y + x) + c] = render[4u * (w * y + x) + c];
		rc = write_pfm(bin, w, h, filename);
	} else {
		unsigned char* bin = (unsigned char*)vis_ptr->binary_buffer;
		for (unsigned int y = 0u; y < h; y++)
			for (unsigned int x = 0u; x < w; x++) {
				f32 color[3];
				unsigned char pixel[3];
				for (int i = 0; i < 3; i++) {
					color[i] = render[4u * (w * y + x) + (unsigned int)i];
					color[i] = 0.0f < color[i] ? color[i] : 0.0f;
					color[i] = 1.0f > color[i] ? color[i] : 1.0f;
				}
				for (int i = 0; i < 3; i++) {
					if (filetype == IMAGE_FILETYPE_PPM)   /* Rec. 709 transfer */
						color[i] = (f32)0.018053968510807 <= color[i]
						               ? (f32)1.099296826809442 * powf(color[i], 0.45f) - (f32)0.099296826809442
						               : 4.500f * color[i];
					else                                  /* sRGB transfer */
						color[i] = (f32)0.0031308 < color[i]
						               ? 1.055f * powf(color[i], (float)(1.0 / 2.4)) - 0.055f
						               : 12.92f * color[i];
				}
				for (int i = 0; i < 3; i++) {
					color[i] *= 255.0f;
					color[i] = roundf(color[i]);
					color[i] = 0.0f < color[i] ? color[i] : 0.0f;
					color[i] = 255.0f > color[i] ? color[i] : 255.0f;
					pixel[i] = (unsigned char)color[i];
				}
				for (int i = 0; i < 3; i++) bin[3u * (w * ((h - 1u) - y) + x) + (unsigned int)i] = pixel[i];
			}
		rc = filetype == IMAGE_FILETYPE_PPM ? write_ppm(bin, w, h, filename) : write_png(bin, w, h, filename);
	}
	if (rc) {
		fprintf(stderr, "%s error: image_write() failed.\n", "vis_save()");
		sf_error("vis_save", "writing \"%s\" failed", filename);
		return STARFLOOD_FAILURE;
	}
	return STARFLOOD_SUCCESS;
}
