/* Seeded 3-D simplex noise with analytic gradient, and multi-octave curl noise.
 *
 * Same interface as the reference's src/noise.h:7-22: it builds the initial flow, so
 * seeded runs start from the same state as the reference's.
 */
#ifndef GGB_NOISE_H
#define GGB_NOISE_H

#include "vec3.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	unsigned char perm[512];
} Noise;

typedef struct {
	double scale;      /* base frequency on the unit sphere */
	double lacunarity; /* frequency multiplier per octave */
	double gain;       /* amplitude multiplier per octave */
	int octaves;
} CurlParams;

void noise_init(Noise *nz, unsigned int seed);
double simplex3(const Noise *nz, vec3 p, vec3 *g);
vec3 curl_noise(const Noise *nz, vec3 p, const CurlParams *cp);

#ifdef __cplusplus
}
#endif

#endif
