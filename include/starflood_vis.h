/*
 * starflood_vis.h -- the reference's visualisation interface
 * (src/visualization.h:10-39): an accumulation rasteriser that splats every
 * particle as SPATIAL_SAMPLES (128) jittered samples into an integer RGBA
 * buffer, a tone-mapped float render buffer, and image files.
 *
 * vis_t has the reference's layout; its buffers are host memory, filled by
 * vis_draw after the splat runs on the device.  The reference's
 * compile-time settings are fixed as in its config.h: SPATIAL_SAMPLES 128,
 * EXPOSURE -1, ORTHO_SCALE -1, AUTOSCALING on, no SHUTTER_SPEED, and the
 * VIS_FILE_FORMAT 2 (PNG) build: tanh tone mapping and an 8-bit binary
 * buffer for PPM and PNG.
 */
#ifndef STARFLOOD_VIS_H
#define STARFLOOD_VIS_H

#include "starflood.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference src/visualization.h:10-16 */
typedef enum {
	IMAGE_FILETYPE_PFM,
	IMAGE_FILETYPE_PPM,
	IMAGE_FILETYPE_PNG,
	NUM_IMAGE_FILETYPE
} image_filetype_t;

/* reference src/visualization.h:18-27 */
typedef struct {
	unsigned int sizex, sizey; /* visualisation dimensions */
	void* mem;                 /* one allocation holding the three buffers */
	i32* atomic_buffer;        /* 4 x sizex x sizey: accumulated R, G, B, sample count */
	f32* render_buffer;        /* 4 x sizex x sizey: tone-mapped RGB, alpha 1 */
	void* binary_buffer;       /* the image file's pixels (12 x sizex x sizey bytes here) */
} vis_t;

/* reference src/visualization.h:30 (src/visualization.c:37-190): allocates
   and zeroes the buffers (4 MiB aligned; page-locked when a device is
   present, so vis_draw copies back at full rate); the device copies are made
   by the first vis_draw. */
int vis_init(vis_t* SF_RESTRICT vis_ptr, const unsigned int sizex, const unsigned int sizey);

/* reference src/visualization.h:33 (src/visualization.c:192-230) */
int vis_free(vis_t* SF_RESTRICT vis_ptr);

/* reference src/visualization.h:36 (src/visualization_io.c:166-330):
   writes the render buffer as PFM (RGB floats), PPM (8-bit, Rec. 709
   transfer) or PNG (8-bit, sRGB transfer; stored deflate blocks).  Rows
   are flipped for PPM and PNG as in the reference.  Fails when the file
   cannot be written (the reference only prints the error). */
int vis_save(const vis_t* SF_RESTRICT vis_ptr, const char* SF_RESTRICT filename, image_filetype_t filetype);

/* reference src/visualization.h:39 (src/visualization.c:253-736): splats
   the simulation's particles (pos, vel, pot, step_number) into the
   accumulation buffer and tone-maps it into the render buffer.  Both are
   copied back into vis_ptr's host buffers. */
int vis_draw(const vis_t* SF_RESTRICT vis_ptr, const sim_t* SF_RESTRICT sim_ptr);

#ifdef __cplusplus
}
#endif

#endif
