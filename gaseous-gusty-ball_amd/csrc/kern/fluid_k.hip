// fluid_k.hip -- gfx950 kernels for the fluid step: Voronoi finite-volume operators,
// the PCG vector passes with device-resident scalars, the velocity reconstruction and
// the BFECC semi-Lagrangian advection.
//
// Every pass is HBM-bound (sparse stencil / gather); reductions are deterministic
// (wave64 tree, then blocks folded in order by the last block to finish), and the
// CG loop's control lives on the device (ggbk_scalars.active), so the host can launch
// iterations ahead without reading scalars back.
//
// T is the state storage type (double, or float in reference-precision mode); per-cell
// arithmetic is double as in the reference (src/fluid.c).
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256

template <typename T>
__device__ __forceinline__ double applyK_at(const ggbk_grid &g, const T *x, int i)
{
	// (K x)_i = sum_k w_ik (x_i - x_nb(k)) in stored neighbour order (fluid.c:277-286)
	double xi = (double) x[i], s = 0;
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++) {
		int j = g.nbr[(size_t) k * g.n + i];
		s += (double) g.w[(size_t) k * g.n + i] * (xi - (double) x[j]);
	}
	return s;
}

// ---------------- free operators ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_curl(ggbk_grid g, const T *vel, T *out)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	dv3 p = cell_pos(g, i);
	dv3 wi = vcross(ld3(vel + 3 * (size_t) i), p);
	double s = 0;
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++) {
		int j = g.nbr[(size_t) k * g.n + i];
		dv3 pj = cell_pos(g, j);
		dv3 wj = vcross(ld3(vel + 3 * (size_t) j), pj);
		dv3 nrm = vnorm(vsub(pj, p));
		s += (double) g.elen[(size_t) k * g.n + i] * 0.5 * vdot(vadd(wi, wj), nrm);
	}
	out[i] = (T) (s / (double) g.area[i]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_divergence(ggbk_grid g, const T *vel, T *out)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	dv3 p = cell_pos(g, i);
	dv3 ui = ld3(vel + 3 * (size_t) i);
	double s = 0;
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++) {
		int j = g.nbr[(size_t) k * g.n + i];
		dv3 nrm = vnorm(vsub(cell_pos(g, j), p));
		dv3 um = vadd(ui, ld3(vel + 3 * (size_t) j));
		s += (double) g.elen[(size_t) k * g.n + i] * 0.5 * vdot(um, nrm);
	}
	out[i] = (T) (s / (double) g.area[i]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_laplacian(ggbk_grid g, const T *x, T *y)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double s = 0, xi = (double) x[i];
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++)
		s += (double) g.w[(size_t) k * g.n + i] * ((double) x[g.nbr[(size_t) k * g.n + i]] - xi);
	y[i] = (T) (s / (double) g.area[i]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_applyK(ggbk_grid g, const T *x, double *y)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		y[i] = applyK_at(g, x, i);
}

// area-weighted sum; the last block stores the mean rounded to T (fluid.c:264-275)
template <typename T>
__global__ __launch_bounds__(BS) void k_mean(ggbk_grid g, const T *f, ggbk_scalars *s, double *part)
{
	int i = blockIdx.x * BS + threadIdx.x;
	double v = i < g.n ? (double) g.area[i] * (double) f[i] : 0.0;
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bsum;
	if (last_block(&s->ticket[1])) {
		double tot = fold_partials<BS>(part, gridDim.x);
		if (threadIdx.x == 0) {
			s->sum = tot;
			s->mean = (double) (T) (tot / g.total_area);
		}
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_sub_mean(int n, T *f, const ggbk_scalars *s)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < n)
		f[i] -= (T) s->mean;
}

// ---------------- PCG (fluid.c:288-358) ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_cg_start(ggbk_grid g, const T *omega, const T *psi, T *r,
						  double tol, int maxit, ggbk_scalars *s, double *part)
{
	int i = blockIdx.x * BS + threadIdx.x;
	double bbl = 0, rrl = 0;
	if (i < g.n) {
		double b = -(double) g.area[i] * (double) omega[i];
		double ri = b - applyK_at(g, psi, i);
		r[i] = (T) ri;
		bbl = b * b;
		rrl = ri * ri;
	}
	double bsum = block_sum<BS>(bbl);
	double rsum = block_sum<BS>(rrl);
	if (threadIdx.x == 0) {
		part[blockIdx.x] = bsum;
		part[gridDim.x + blockIdx.x] = rsum;
	}
	if (last_block(&s->ticket[0])) {
		double bb = fold_partials<BS>(part, gridDim.x);
		double rr = fold_partials<BS>(part + gridDim.x, gridDim.x);
		if (threadIdx.x == 0) {
			s->bb = bb;
			s->rr = rr;
			s->it = 0;
			s->maxit = maxit;
			s->rz = 0;
			if (bb == 0.0) {
				s->status = 2;
				s->tol2 = 0;
				s->active = 0;
			} else {
				s->status = 0;
				s->tol2 = tol * tol * bb;
				s->active = (rr > s->tol2 && maxit > 0) ? 1 : 0;
			}
		}
	}
}

template <typename T, typename TZ>
__global__ __launch_bounds__(BS) void k_cg_first_dir(int n, const T *r, const TZ *z, T *p, ggbk_scalars *s,
						      double *part)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	double v = 0;
	if (i < n) {
		p[i] = (T) z[i];
		v = (double) r[i] * (double) z[i];
	}
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bsum;
	if (last_block(&s->ticket[2])) {
		double rz = fold_partials<BS>(part, gridDim.x);
		if (threadIdx.x == 0)
			s->rz = rz;
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_cg_Ap(ggbk_grid g, const T *p, T *Ap, ggbk_scalars *s, double *part)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	double v = 0;
	if (i < g.n) {
		double a = applyK_at(g, p, i);
		Ap[i] = (T) a;
		v = a * (double) p[i];
	}
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bsum;
	if (last_block(&s->ticket[3])) {
		double pAp = fold_partials<BS>(part, gridDim.x);
		if (threadIdx.x == 0) {
			s->pAp = pAp;
			if (pAp <= 0) {
				s->status = 1;
				s->active = 0;
			} else {
				s->alpha = s->rz / pAp;
			}
		}
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_cg_update(int n, T *psi, T *r, const T *p, const T *Ap, ggbk_scalars *s,
						   double *part)
{
	if (!s->active)
		return;
	double alpha = s->alpha;
	int i = blockIdx.x * BS + threadIdx.x;
	double v = 0;
	if (i < n) {
		psi[i] += (T) (alpha * (double) p[i]);
		double ri = (double) r[i] - alpha * (double) Ap[i];
		r[i] = (T) ri;
		v = ri * ri;
	}
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bsum;
	if (last_block(&s->ticket[4])) {
		double rr = fold_partials<BS>(part, gridDim.x);
		if (threadIdx.x == 0) {
			s->rr = rr;
			int it = s->it + 1;
			s->it = it;
			if (rr <= s->tol2) {
				s->active = 0;
				s->status = 0;
			} else if (it >= s->maxit) {
				s->active = 0;
				s->status = 3;
			}
		}
	}
}

template <typename T, typename TZ>
__global__ __launch_bounds__(BS) void k_cg_rz(int n, const T *r, const TZ *z, ggbk_scalars *s, double *part)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	double v = i < n ? (double) r[i] * (double) z[i] : 0.0;
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bsum;
	if (last_block(&s->ticket[5])) {
		double rznew = fold_partials<BS>(part, gridDim.x);
		if (threadIdx.x == 0) {
			s->beta = rznew / s->rz;
			s->rz = rznew;
		}
	}
}

template <typename T, typename TZ>
__global__ __launch_bounds__(BS) void k_cg_pupdate(int n, const TZ *z, T *p, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	double beta = s->beta;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < n)
		p[i] = (T) ((double) z[i] + beta * (double) p[i]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_jacobi_precond(ggbk_grid g, const T *r, T *z, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		z[i] = g.dinv[i] * r[i];
}

template <typename T>
__global__ __launch_bounds__(BS) void k_zero_if_bb0(int n, T *psi, const ggbk_scalars *s)
{
	if (s->status != 2)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < n)
		psi[i] = (T) 0;
}

// ---------------- velocity and max speed ----------------

// u = p x sum_k gc_k (psi_j - psi_i) (fluid.c:360-376); max |u|^2 of the stored values
// in storage arithmetic (fluid.c:480-491) folded into s->vmax2.
template <typename T>
__global__ __launch_bounds__(BS) void k_velocity(ggbk_grid g, const T *psi, T *vel, ggbk_scalars *s, double *part)
{
	int i = blockIdx.x * BS + threadIdx.x;
	double m = 0;
	if (i < g.n) {
		dv3 p = cell_pos(g, i);
		dv3 gr = d3(0, 0, 0);
		double pi = (double) psi[i];
		int nn = g.nnb[i];
		for (int k = 0; k < nn; k++) {
			int j = g.nbr[(size_t) k * g.n + i];
			dv3 c = d3(g.gc[((size_t) k * 3 + 0) * g.n + i], g.gc[((size_t) k * 3 + 1) * g.n + i],
				   g.gc[((size_t) k * 3 + 2) * g.n + i]);
			gr = vadd(gr, vscl(c, (double) psi[j] - pi));
		}
		dv3 u = vcross(p, gr);
		T u0 = (T) u.x, u1 = (T) u.y, u2 = (T) u.z;
		vel[3 * (size_t) i + 0] = u0;
		vel[3 * (size_t) i + 1] = u1;
		vel[3 * (size_t) i + 2] = u2;
		T sq = u0 * u0 + u1 * u1 + u2 * u2;
		m = (double) sq;
	}
	double bm = block_max<BS>(m);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bm;
	if (last_block(&s->ticket[6])) {
		double v = fold_partials_max<BS>(part, gridDim.x);
		if (threadIdx.x == 0)
			s->vmax2 = v;
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_max_speed2(int n, const T *vel, ggbk_scalars *s, double *part)
{
	int i = blockIdx.x * BS + threadIdx.x;
	double m = 0;
	if (i < n) {
		const T *v = vel + 3 * (size_t) i;
		T sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
		m = (double) sq;
	}
	double bm = block_max<BS>(m);
	if (threadIdx.x == 0)
		part[blockIdx.x] = bm;
	if (last_block(&s->ticket[6])) {
		double v = fold_partials_max<BS>(part, gridDim.x);
		if (threadIdx.x == 0)
			s->vmax2 = v;
	}
}

// ---------------- semi-Lagrangian advection (fluid.c:397-478) ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_departures(ggbk_grid g, const T *vel, T *dep_b, T *dep_f, double dt, int fwd)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	dv3 p = cell_pos(g, i);
	dv3 u = ld3(vel + 3 * (size_t) i);
	dv3 pm = gc_move(p, u, -0.5 * dt);
	dv3 um = vtan(sample_vec(g, vel, pm), p);
	st3(dep_b + 3 * (size_t) i, gc_move(p, um, -dt));
	if (fwd) {
		pm = gc_move(p, u, 0.5 * dt);
		um = vtan(sample_vec(g, vel, pm), p);
		st3(dep_f + 3 * (size_t) i, gc_move(p, um, dt));
	}
}

template <typename T>
__global__ __launch_bounds__(BS) void k_absvort(ggbk_grid g, const T *omega, T *q, double twoOmega)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		q[i] = (T) ((double) omega[i] + twoOmega * (double) g.pos[3 * (size_t) i + 1]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_advect(ggbk_grid g, const T *src, const T *dep, T *dst)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		dst[i] = (T) sample_scalar(g, src, ld3(dep + 3 * (size_t) i));
}

template <typename T>
__global__ __launch_bounds__(BS) void k_bfecc_source(int n, const T *q, T *q2)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < n)
		q2[i] = q[i] + (T) 0.5f * (q[i] - q2[i]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_limited_advect(ggbk_grid g, const T *q2, const T *q, const T *dep_b, T *q1)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	Taps t = taps_at(g, ld3(dep_b + 3 * (size_t) i));
	double wt[4];
	tap_weights(t, wt);
	double v = 0, lo = 1e30, hi = -1e30;
	for (int k = 0; k < 4; k++) {
		v += wt[k] * (double) q2[t.idx[k]];
		double o = (double) q[t.idx[k]];
		if (o < lo)
			lo = o;
		if (o > hi)
			hi = o;
	}
	if (v < lo)
		v = lo;
	if (v > hi)
		v = hi;
	q1[i] = (T) v;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_relax(ggbk_grid g, const T *qn, const T *band, T *omega, double twoOmega,
					       double relax)
{
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double w = (double) qn[i] - twoOmega * (double) g.pos[3 * (size_t) i + 1];
	w += relax * ((double) band[i] - w);
	omega[i] = (T) w;
}

// ---------------- launchers ----------------

#define NB(n) ggbk_blocks((n), BS)

extern "C" void ggbk_curl(ggbk_grid g, const void *vel, void *out, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_curl<double>, NB(g.n), BS, 0, g, (const double *) vel, (double *) out);
	else
		GGBK_LAUNCH(k_curl<float>, NB(g.n), BS, 0, g, (const float *) vel, (float *) out);
}

extern "C" void ggbk_divergence(ggbk_grid g, const void *vel, void *out, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_divergence<double>, NB(g.n), BS, 0, g, (const double *) vel, (double *) out);
	else
		GGBK_LAUNCH(k_divergence<float>, NB(g.n), BS, 0, g, (const float *) vel, (float *) out);
}

extern "C" void ggbk_laplacian(ggbk_grid g, const void *x, void *y, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_laplacian<double>, NB(g.n), BS, 0, g, (const double *) x, (double *) y);
	else
		GGBK_LAUNCH(k_laplacian<float>, NB(g.n), BS, 0, g, (const float *) x, (float *) y);
}

extern "C" void ggbk_applyK(ggbk_grid g, const void *x, double *y, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_applyK<double>, NB(g.n), BS, 0, g, (const double *) x, y);
	else
		GGBK_LAUNCH(k_applyK<float>, NB(g.n), BS, 0, g, (const float *) x, y);
}

extern "C" void ggbk_remove_mean(ggbk_grid g, void *f, int T, ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64) {
		GGBK_LAUNCH(k_mean<double>, NB(g.n), BS, 0, g, (const double *) f, s, part);
		GGBK_LAUNCH(k_sub_mean<double>, NB(g.n), BS, 0, g.n, (double *) f, (const ggbk_scalars *) s);
	} else {
		GGBK_LAUNCH(k_mean<float>, NB(g.n), BS, 0, g, (const float *) f, s, part);
		GGBK_LAUNCH(k_sub_mean<float>, NB(g.n), BS, 0, g.n, (float *) f, (const ggbk_scalars *) s);
	}
}

extern "C" void ggbk_cg_start(ggbk_grid g, const void *omega, const void *psi, void *r, int T, double tol,
			      int maxit, ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_cg_start<double>, NB(g.n), BS, 0, g, (const double *) omega, (const double *) psi,
			    (double *) r, tol, maxit, s, part);
	else
		GGBK_LAUNCH(k_cg_start<float>, NB(g.n), BS, 0, g, (const float *) omega, (const float *) psi,
			    (float *) r, tol, maxit, s, part);
}

extern "C" void ggbk_cg_first_dir(ggbk_grid g, const void *r, const void *z, int TZ, void *p, int T,
				  ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64 && TZ == GGBK_F64)
		GGBK_LAUNCH((k_cg_first_dir<double, double>), NB(g.n), BS, 0, g.n, (const double *) r,
			    (const double *) z, (double *) p, s, part);
	else if (T == GGBK_F64)
		GGBK_LAUNCH((k_cg_first_dir<double, float>), NB(g.n), BS, 0, g.n, (const double *) r,
			    (const float *) z, (double *) p, s, part);
	else
		GGBK_LAUNCH((k_cg_first_dir<float, float>), NB(g.n), BS, 0, g.n, (const float *) r,
			    (const float *) z, (float *) p, s, part);
}

extern "C" void ggbk_cg_Ap(ggbk_grid g, const void *p, void *Ap, int T, ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_cg_Ap<double>, NB(g.n), BS, 0, g, (const double *) p, (double *) Ap, s, part);
	else
		GGBK_LAUNCH(k_cg_Ap<float>, NB(g.n), BS, 0, g, (const float *) p, (float *) Ap, s, part);
}

extern "C" void ggbk_cg_update(ggbk_grid g, void *psi, void *r, const void *p, const void *Ap, int T,
			       ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_cg_update<double>, NB(g.n), BS, 0, g.n, (double *) psi, (double *) r,
			    (const double *) p, (const double *) Ap, s, part);
	else
		GGBK_LAUNCH(k_cg_update<float>, NB(g.n), BS, 0, g.n, (float *) psi, (float *) r, (const float *) p,
			    (const float *) Ap, s, part);
}

extern "C" void ggbk_cg_next_dir(ggbk_grid g, const void *r, const void *z, int TZ, void *p, int T,
				 ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64 && TZ == GGBK_F64) {
		GGBK_LAUNCH((k_cg_rz<double, double>), NB(g.n), BS, 0, g.n, (const double *) r, (const double *) z,
			    s, part);
		GGBK_LAUNCH((k_cg_pupdate<double, double>), NB(g.n), BS, 0, g.n, (const double *) z, (double *) p,
			    (const ggbk_scalars *) s);
	} else if (T == GGBK_F64) {
		GGBK_LAUNCH((k_cg_rz<double, float>), NB(g.n), BS, 0, g.n, (const double *) r, (const float *) z, s,
			    part);
		GGBK_LAUNCH((k_cg_pupdate<double, float>), NB(g.n), BS, 0, g.n, (const float *) z, (double *) p,
			    (const ggbk_scalars *) s);
	} else {
		GGBK_LAUNCH((k_cg_rz<float, float>), NB(g.n), BS, 0, g.n, (const float *) r, (const float *) z, s,
			    part);
		GGBK_LAUNCH((k_cg_pupdate<float, float>), NB(g.n), BS, 0, g.n, (const float *) z, (float *) p,
			    (const ggbk_scalars *) s);
	}
}

extern "C" void ggbk_jacobi_precond(ggbk_grid g, const void *r, void *z, int T, const ggbk_scalars *s)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_jacobi_precond<double>, NB(g.n), BS, 0, g, (const double *) r, (double *) z, s);
	else
		GGBK_LAUNCH(k_jacobi_precond<float>, NB(g.n), BS, 0, g, (const float *) r, (float *) z, s);
}

extern "C" void ggbk_cg_zero_if_bb0(ggbk_grid g, void *psi, int T, const ggbk_scalars *s)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_zero_if_bb0<double>, NB(g.n), BS, 0, g.n, (double *) psi, s);
	else
		GGBK_LAUNCH(k_zero_if_bb0<float>, NB(g.n), BS, 0, g.n, (float *) psi, s);
}

extern "C" void ggbk_velocity(ggbk_grid g, const void *psi, void *vel, int T, ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_velocity<double>, NB(g.n), BS, 0, g, (const double *) psi, (double *) vel, s, part);
	else
		GGBK_LAUNCH(k_velocity<float>, NB(g.n), BS, 0, g, (const float *) psi, (float *) vel, s, part);
}

extern "C" void ggbk_max_speed2(ggbk_grid g, const void *vel, int T, ggbk_scalars *s, double *part)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_max_speed2<double>, NB(g.n), BS, 0, g.n, (const double *) vel, s, part);
	else
		GGBK_LAUNCH(k_max_speed2<float>, NB(g.n), BS, 0, g.n, (const float *) vel, s, part);
}

extern "C" void ggbk_departures(ggbk_grid g, const void *vel, void *dep_b, void *dep_f, int T, double dt, int fwd)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_departures<double>, NB(g.n), BS, 0, g, (const double *) vel, (double *) dep_b,
			    (double *) dep_f, dt, fwd);
	else
		GGBK_LAUNCH(k_departures<float>, NB(g.n), BS, 0, g, (const float *) vel, (float *) dep_b,
			    (float *) dep_f, dt, fwd);
}

extern "C" void ggbk_absvort(ggbk_grid g, const void *omega, void *q, int T, double twoOmega)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_absvort<double>, NB(g.n), BS, 0, g, (const double *) omega, (double *) q, twoOmega);
	else
		GGBK_LAUNCH(k_absvort<float>, NB(g.n), BS, 0, g, (const float *) omega, (float *) q, twoOmega);
}

extern "C" void ggbk_advect(ggbk_grid g, const void *src, const void *dep, void *dst, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_advect<double>, NB(g.n), BS, 0, g, (const double *) src, (const double *) dep,
			    (double *) dst);
	else
		GGBK_LAUNCH(k_advect<float>, NB(g.n), BS, 0, g, (const float *) src, (const float *) dep,
			    (float *) dst);
}

extern "C" void ggbk_bfecc_source(int n, const void *q, void *q2, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_bfecc_source<double>, NB(n), BS, 0, n, (const double *) q, (double *) q2);
	else
		GGBK_LAUNCH(k_bfecc_source<float>, NB(n), BS, 0, n, (const float *) q, (float *) q2);
}

extern "C" void ggbk_limited_advect(ggbk_grid g, const void *q2, const void *q, const void *dep_b, void *q1, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_limited_advect<double>, NB(g.n), BS, 0, g, (const double *) q2, (const double *) q,
			    (const double *) dep_b, (double *) q1);
	else
		GGBK_LAUNCH(k_limited_advect<float>, NB(g.n), BS, 0, g, (const float *) q2, (const float *) q,
			    (const float *) dep_b, (float *) q1);
}

extern "C" void ggbk_relax(ggbk_grid g, const void *qn, const void *band, void *omega, int T, double twoOmega,
			   double relax)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_relax<double>, NB(g.n), BS, 0, g, (const double *) qn, (const double *) band,
			    (double *) omega, twoOmega, relax);
	else
		GGBK_LAUNCH(k_relax<float>, NB(g.n), BS, 0, g, (const float *) qn, (const float *) band,
			    (float *) omega, twoOmega, relax);
}
