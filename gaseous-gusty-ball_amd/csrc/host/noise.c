/*
 * Seeded 3-D simplex noise with analytic gradient and fractal curl noise (host).
 *
 * Builds the initial flow once per run.  Seeded runs must start from the reference's
 * state, so the permutation generator (xorshift64* Fisher-Yates from a seed mix,
 * reference noise.c:12-31), the simplex skew/unskew, corner ordering, hash and
 * gradient table (noise.c:33-79), the 72x scale, and the curl-of-three-potentials
 * octave sum with its fixed potential and octave offsets (noise.c:81-103) are the
 * reference's; only the code is this library's.
 */
#include <math.h>
#include <stdint.h>

#include "noise.h"

/* the 12 edge-midpoint gradients of the cube */
static const signed char GRAD[12][3] = {
	{ 1, 1, 0 }, { -1, 1, 0 }, { 1, -1, 0 }, { -1, -1, 0 },
	{ 1, 0, 1 }, { -1, 0, 1 }, { 1, 0, -1 }, { -1, 0, -1 },
	{ 0, 1, 1 }, { 0, -1, 1 }, { 0, 1, -1 }, { 0, -1, -1 },
};

void noise_init(Noise *nz, unsigned int seed)
{
	unsigned char p[256];
	uint64_t s = 0x9E3779B97F4A7C15ULL ^ ((uint64_t) seed * 0xD1B54A32D192ED03ULL);
	for (int i = 0; i < 256; i++)
		p[i] = (unsigned char) i;
	for (int i = 255; i > 0; i--) {
		s ^= s >> 12;
		s ^= s << 25;
		s ^= s >> 27;
		uint64_t draw = s * 2685821657736338717ULL;
		int j = (int) ((draw >> 33) % (uint64_t) (i + 1));
		unsigned char t = p[i];
		p[i] = p[j];
		p[j] = t;
	}
	for (int i = 0; i < 512; i++)
		nz->perm[i] = p[i & 255];
}

double simplex3(const Noise *nz, vec3 pt, vec3 *grad)
{
	const double F3 = 1.0 / 3.0, G3 = 1.0 / 6.0;
	const unsigned char *pm = nz->perm;
	double skew = (pt.x + pt.y + pt.z) * F3;
	int i = (int) floor(pt.x + skew), j = (int) floor(pt.y + skew), k = (int) floor(pt.z + skew);
	double unskew = (i + j + k) * G3;
	double x0 = pt.x - (i - unskew), y0 = pt.y - (j - unskew), z0 = pt.z - (k - unskew);
	/* second and third corners of the simplex containing the point */
	int o1[3], o2[3];
	if (x0 >= y0) {
		if (y0 >= z0) {
			o1[0] = 1; o1[1] = 0; o1[2] = 0; o2[0] = 1; o2[1] = 1; o2[2] = 0;
		} else if (x0 >= z0) {
			o1[0] = 1; o1[1] = 0; o1[2] = 0; o2[0] = 1; o2[1] = 0; o2[2] = 1;
		} else {
			o1[0] = 0; o1[1] = 0; o1[2] = 1; o2[0] = 1; o2[1] = 0; o2[2] = 1;
		}
	} else {
		if (y0 < z0) {
			o1[0] = 0; o1[1] = 0; o1[2] = 1; o2[0] = 0; o2[1] = 1; o2[2] = 1;
		} else if (x0 < z0) {
			o1[0] = 0; o1[1] = 1; o1[2] = 0; o2[0] = 0; o2[1] = 1; o2[2] = 1;
		} else {
			o1[0] = 0; o1[1] = 1; o1[2] = 0; o2[0] = 1; o2[1] = 1; o2[2] = 0;
		}
	}
	double cx[4] = { x0, x0 - o1[0] + G3, x0 - o2[0] + 2 * G3, x0 - 1 + 3 * G3 };
	double cy[4] = { y0, y0 - o1[1] + G3, y0 - o2[1] + 2 * G3, y0 - 1 + 3 * G3 };
	double cz[4] = { z0, z0 - o1[2] + G3, z0 - o2[2] + 2 * G3, z0 - 1 + 3 * G3 };
	int ii = i & 255, jj = j & 255, kk = k & 255;
	int hash[4];
	hash[0] = pm[ii + pm[jj + pm[kk]]] % 12;
	hash[1] = pm[ii + o1[0] + pm[jj + o1[1] + pm[kk + o1[2]]]] % 12;
	hash[2] = pm[ii + o2[0] + pm[jj + o2[1] + pm[kk + o2[2]]]] % 12;
	hash[3] = pm[ii + 1 + pm[jj + 1 + pm[kk + 1]]] % 12;

	double val = 0, gx = 0, gy = 0, gz = 0;
	for (int c = 0; c < 4; c++) {
		double t = 0.5 - cx[c] * cx[c] - cy[c] * cy[c] - cz[c] * cz[c];
		if (t <= 0)
			continue;
		const signed char *gr = GRAD[hash[c]];
		double gd = gr[0] * cx[c] + gr[1] * cy[c] + gr[2] * cz[c];
		double t2 = t * t, t4 = t2 * t2, t3 = t2 * t;
		val += t4 * gd;
		gx += t4 * gr[0] - 8.0 * t3 * gd * cx[c];
		gy += t4 * gr[1] - 8.0 * t3 * gd * cy[c];
		gz += t4 * gr[2] - 8.0 * t3 * gd * cz[c];
	}
	if (grad)
		*grad = v3(gx * 72.0, gy * 72.0, gz * 72.0);
	return val * 72.0;
}

vec3 curl_noise(const Noise *nz, vec3 p, const CurlParams *cp)
{
	/* offsets that decorrelate the three potential components */
	static const double POT[3][3] = { { 0.0, 0.0, 0.0 }, { 31.416, -47.853, 12.793 }, { -93.11, 63.2, 109.7 } };
	vec3 acc = v3(0, 0, 0);
	double freq = cp->scale, amp = 1.0;
	for (int o = 0; o < cp->octaves; o++) {
		vec3 q = v3_add(v3_scale(p, freq), v3(17.13 * o, 5.31 * o, 11.71 * o));
		vec3 g[3];
		for (int c = 0; c < 3; c++)
			simplex3(nz, v3_add(q, v3(POT[c][0], POT[c][1], POT[c][2])), &g[c]);
		double m = amp * freq;
		vec3 curl = v3(g[2].y - g[1].z, g[0].z - g[2].x, g[1].x - g[0].y);
		acc = v3_add(acc, v3_scale(curl, m));
		freq *= cp->lacunarity;
		amp *= cp->gain;
	}
	return acc;
}
