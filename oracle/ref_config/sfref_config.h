/*
 * sfref_config.h -- configuration of the reference build in oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference is configured by editing its
 * src/config.h; this header includes that file unchanged (found through
 * -I<reference>/src) and then overrides the macros a configuration needs,
 * the way a user edits config.h.  It is prepended to each reference source
 * with gcc -include, so config.h's #pragma once makes the sources' own
 * #include "config.h" a no-op.
 *
 * Set per build with -D:
 *   SFREF_EPSILON, SFREF_DT        softening and time step
 *   SFREF_TREE                     keep ENABLE_TREE (else direct summation)
 *   SFREF_SPH                      define ENABLE_SPH
 *   SFREF_TREE_NODES_MAX           override TREE_NODES_MAX
 *   SFREF_NUM_BODIES, SFREF_NUM_TIMESTEPS, SFREF_OUTPUT_INTERVAL
 *                                  the run of src/main.c (the program
 *                                  builds in _ref/ref_main_*)
 */
#include "config.h"

/* debug overlay and offload off; no per-phase printing */
#undef SOLVER_VIS
#undef ENABLE_OFFLOAD_SIM
#undef ENABLE_OFFLOAD_VIS
#undef ENABLE_TIMING
#undef ENABLE_VIS
/* the visualisation is built without its stb PNG writer (absent here; PFM
   and PPM stay) and without its timing log */
#undef ENABLE_STB_IMAGE_WRITE
#undef LOG_TIMINGS_VIS_DRAW

#undef EPSILON
#define EPSILON SFREF_EPSILON
#undef TIMESTEP_SIZE
#define TIMESTEP_SIZE SFREF_DT

#ifndef SFREF_TREE
#undef ENABLE_TREE
#endif

#ifdef SFREF_SPH
#define ENABLE_SPH
#endif

#ifdef SFREF_TREE_NODES_MAX
#undef TREE_NODES_MAX
#define TREE_NODES_MAX SFREF_TREE_NODES_MAX
#endif

#ifdef SFREF_NUM_BODIES
#undef NUM_BODIES
#define NUM_BODIES SFREF_NUM_BODIES
#endif

#ifdef SFREF_NUM_TIMESTEPS
#undef NUM_TIMESTEPS
#define NUM_TIMESTEPS SFREF_NUM_TIMESTEPS
#endif

#ifdef SFREF_OUTPUT_INTERVAL
#undef OUTPUT_INTERVAL
#define OUTPUT_INTERVAL SFREF_OUTPUT_INTERVAL
#endif
