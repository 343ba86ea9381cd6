// particles_k.hip -- tracer particles on the device: Fibonacci-lattice start
// positions, equirectangular colour lookup, and RK2 midpoint advection along great
// circles through the bilinear velocity field (reference src/particles.c:8-74).
//
// Positions are stored in double.  With round32 set, every stored position is rounded
// to float, which reproduces the reference's float particle arrays.  The kernels work
// per storage slot; the slot order (pid) only matters to the host's index-ordered views.
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256

__device__ __forceinline__ void store_pos(double *p, dv3 a, int round32)
{
	if (round32)
		a = d3((double) (float) a.x, (double) (float) a.y, (double) (float) a.z);
	st3(p, a);
}

__global__ __launch_bounds__(BS) void k_particles_init(ggbk_parts ps, int round32)
{
	int i = xblock() * BS + threadIdx.x;
	int n = ps.n;
	if (i >= n)
		return;
	const double golden = GGB_PI * (3.0 - sqrt(5.0));
	double y = 1.0 - 2.0 * (i + 0.5) / n;
	double r = sqrt(1.0 - y * y);
	double phi = golden * i;
	dv3 p = d3(r * cos(phi), y, r * sin(phi));
	store_pos(ps.pos + 3 * (size_t) i, p, round32);
	store_pos(ps.prev + 3 * (size_t) i, p, round32);
	for (int c = 0; c < 4; c++)
		ps.col[4 * (size_t) i + c] = 1.0f;
	ps.pid[i] = i;
}

// bilinear colour at the particle's longitude/latitude; x wraps, y clamps
__global__ __launch_bounds__(BS) void k_particles_color(int n, const double *pos, float *col, const float *img, int w,
							int h)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= n)
		return;
	dv3 d = ld3(pos + 3 * (size_t) i);
	double lon = atan2(d.x, d.z);
	double yy = d.y > 1.0 ? 1.0 : (d.y < -1.0 ? -1.0 : d.y);
	double lat = asin(yy);
	double x = (lon + GGB_PI) / (2 * GGB_PI) * w - 0.5;
	double y = (GGB_PI / 2 - lat) / GGB_PI * h - 0.5;
	int x0 = (int) floor(x), y0 = (int) floor(y);
	double fx = x - x0, fy = y - y0;
	int xs[2] = { ((x0 % w) + w) % w, (((x0 + 1) % w) + w) % w };
	int ys[2] = { y0 < 0 ? 0 : (y0 >= h ? h - 1 : y0), y0 + 1 < 0 ? 0 : (y0 + 1 >= h ? h - 1 : y0 + 1) };
	double wts[4] = { (1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy };
	for (int c = 0; c < 4; c++) {
		double v = 0;
		for (int k = 0; k < 4; k++)
			v += wts[k] * (double) img[((size_t) ys[k >> 1] * w + xs[k & 1]) * 4 + c];
		col[4 * (size_t) i + c] = (float) v;
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_particles_advect(ggbk_grid g, int np, double *pos, double *prev, const T *vel,
							 double dt, int round32)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= np)
		return;
	dv3 p = ld3(pos + 3 * (size_t) i);
	st3(prev + 3 * (size_t) i, p);
	dv3 u1 = sample_vec(g, vel, p);
	dv3 pm = gc_move(p, u1, 0.5 * dt);
	dv3 u2 = vtan(sample_vec(g, vel, pm), p);
	store_pos(pos + 3 * (size_t) i, gc_move(p, u2, dt), round32);
}

extern "C" void ggbk_particles_init(ggbk_parts p, int round32)
{
	GGBK_LAUNCH(k_particles_init, ggbk_blocks(p.n, BS), BS, 0, p, round32);
}

extern "C" void ggbk_particles_color(int n, const double *pos, float *col, const float *img, int w, int h)
{
	GGBK_LAUNCH(k_particles_color, ggbk_blocks(n, BS), BS, 0, n, pos, col, img, w, h);
}

extern "C" void ggbk_particles_advect(ggbk_grid g, int np, double *pos, double *prev, const void *vel, int T,
				      double dt, int round32)
{
	ggbk_probe_begin(GGBK_PROBE_PARTICLES);
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_particles_advect<double>, ggbk_blocks(np, BS), BS, 0, g, np, pos, prev,
			    (const double *) vel, dt, round32);
	else
		GGBK_LAUNCH(k_particles_advect<float>, ggbk_blocks(np, BS), BS, 0, g, np, pos, prev,
			    (const float *) vel, dt, round32);
	ggbk_probe_end(GGBK_PROBE_PARTICLES);
}
