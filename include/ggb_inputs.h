/* Seeded synthetic inputs (host).  No counterpart in the reference headers: the
 * initial flow is the reference driver's static initial_velocity (main.c:320-360),
 * exported here so any caller can build the same initial state; the banded strip is
 * the benchmark's input image. */
#ifndef GGB_INPUTS_H
#define GGB_INPUTS_H

#include <stddef.h>

#include "cubesphere.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Curl noise (tangent, RMS-normalised to noise_amp) plus `bands` zonal bands of speed
 * band_amp; writes vel and the band-only field vband ([ncells*3] each).  0 on success. */
int ggb_initial_velocity(const CubeSphere *cs, unsigned int seed, int bands, double band_amp,
			 double noise_amp, double noise_scale, double lacunarity, double gain,
			 int octaves, float *vel, float *vband);
/* W x H RGBA8 strip of nbands horizontal bands (heights in [hmin, hmax], colours uniform,
 * alpha 255) from a splitmix64 stream seeded by seed.  0 on success. */
int ggb_banded_strip(unsigned int seed, int w, int h, int nbands, int hmin, int hmax,
		     unsigned char *rgba);
/* RGBA8 -> float RGBA in [0, 1] (x / 255). */
void ggb_rgba8_to_float(const unsigned char *rgba, size_t npix, float *out);

#ifdef __cplusplus
}
#endif

#endif
