/*
 * rng.h -- pcg4d, the 4-D permuted-congruential hash of Jarzynski and Olano,
 * "Hash Functions for GPU Rendering", JCGT 9(3), 2020 (the generator the
 * reference uses, src/rng.h:13-40): one LCG step per lane, two rounds of
 * cross-lane multiply-adds around a 16-bit xorshift.
 */
#ifndef SF_RNG_H
#define SF_RNG_H

#include <stdint.h>

/* 2^-32 rounded down to a binary32 value, so u * SF_INV_U32 < 1 (src/rng.h:9) */
#define SF_INV_U32 2.3283062977608182109179324470460414886474609375e-10

#ifdef __HIPCC__
#define SF_RNG_FN static inline __host__ __device__
#else
#define SF_RNG_FN static inline
#endif

SF_RNG_FN void sf_pcg4d(uint32_t s[4]) {
	uint32_t x = s[0] * 1664525u + 1013904223u;
	uint32_t y = s[1] * 1664525u + 1013904223u;
	uint32_t z = s[2] * 1664525u + 1013904223u;
	uint32_t w = s[3] * 1664525u + 1013904223u;
	x += y * w;
	y += z * x;
	z += x * y;
	w += y * z;
	x ^= x >> 16;
	y ^= y >> 16;
	z ^= z >> 16;
	w ^= w >> 16;
	x += y * w;
	y += z * x;
	z += x * y;
	w += y * z;
	s[0] = x;
	s[1] = y;
	s[2] = z;
	s[3] = w;
}

#endif
