/*
 * Seeded synthetic inputs.
 *
 * ggb_initial_velocity: the reference driver's initial flow (main.c:320-360):
 * tangent-projected curl noise normalised to an area-weighted RMS speed, plus zonal
 * bands u_east = A sin(n colatitude).  Sums run in cell order (deterministic).
 *
 * ggb_banded_strip: the benchmark's input image -- a W x H RGBA8 strip of horizontal
 * bands, band heights uniform in [hmin, hmax] rows and colours uniform in [0, 255]^3,
 * alpha 255, drawn from a splitmix64 stream seeded by `seed`.  Bands fill the strip
 * from the top; the last band is extended (or cut) to end at the bottom row.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "cubesphere.h"
#include "ggb_inputs.h"
#include "noise.h"

int ggb_initial_velocity(const CubeSphere *cs, unsigned int seed, int bands, double band_amp, double noise_amp,
			 double noise_scale, double lacunarity, double gain, int octaves, float *vel, float *vband)
{
	int n = cs->ncells;
	Noise nz;
	noise_init(&nz, seed);
	CurlParams cp = { noise_scale, lacunarity, gain, octaves };
	double *term = malloc(sizeof(double) * (size_t) n);
	if (!term)
		return -1;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		vec3 p = cs_pos(cs, i);
		vec3 v = octaves > 0 ? v3_tangent(curl_noise(&nz, p, &cp), p) : v3(0, 0, 0);
		v3_store(vel + 3 * i, v);
		term[i] = v3_dot(v, v) * cs->area[i];
	}
	double ss = 0;
	for (int i = 0; i < n; i++)
		ss += term[i];
	free(term);
	double rms = sqrt(ss / cs->total_area);
	double sc = rms > 0 ? noise_amp / rms : 0;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		vec3 p = cs_pos(cs, i);
		double colat = acos(p.y > 1 ? 1 : (p.y < -1 ? -1 : p.y));
		double s = sin(colat);
		double mag = band_amp * sin(bands * colat);
		vec3 east = v3_cross(v3(0, 1, 0), p);
		vec3 b = s > 1e-9 ? v3_scale(east, mag / s) : v3(0, 0, 0);
		v3_store(vband + 3 * i, b);
		v3_store(vel + 3 * i, v3_add(v3_scale(v3_load(vel + 3 * i), sc), b));
	}
	if (octaves > 0) {
		double fmax = noise_scale * pow(lacunarity, octaves - 1);
		if (1.0 / fmax < 4 * cs->h)
			fprintf(stderr,
				"warning: finest noise octave (feature size %.4f) is smaller than ~4 fluid cells "
				"(%.4f); increase --grid or reduce --noise-octaves\n",
				1.0 / fmax, 4 * cs->h);
	}
	return 0;
}

static uint64_t splitmix64(uint64_t *s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

/* uniform integer in [lo, hi] */
static int uniform_int(uint64_t *s, int lo, int hi)
{
	return lo + (int) (splitmix64(s) % (uint64_t) (hi - lo + 1));
}

int ggb_banded_strip(unsigned int seed, int w, int h, int nbands, int hmin, int hmax, unsigned char *rgba)
{
	if (w < 1 || h < 1 || nbands < 1 || hmin < 1 || hmax < hmin)
		return -1;
	uint64_t s = (uint64_t) seed;
	int row = 0;
	for (int b = 0; b < nbands && row < h; b++) {
		int bh = uniform_int(&s, hmin, hmax);
		int c0 = uniform_int(&s, 0, 255), c1 = uniform_int(&s, 0, 255), c2 = uniform_int(&s, 0, 255);
		int end = b == nbands - 1 ? h : row + bh;
		if (end > h)
			end = h;
		for (; row < end; row++)
			for (int x = 0; x < w; x++) {
				unsigned char *px = rgba + ((size_t) row * w + x) * 4;
				px[0] = (unsigned char) c0;
				px[1] = (unsigned char) c1;
				px[2] = (unsigned char) c2;
				px[3] = 255;
			}
	}
	return 0;
}

/* RGBA8 -> float RGBA in [0, 1] (the reference loader's output format, image_io.c:8-34) */
void ggb_rgba8_to_float(const unsigned char *rgba, size_t npix, float *out)
{
	for (size_t i = 0; i < 4 * npix; i++)
		out[i] = rgba[i] / 255.0f;
}
