// fluid_k.hip -- gfx950 kernels for the fluid step: Voronoi finite-volume operators,
// the PCG vector passes with device-resident scalars, the velocity reconstruction and
// the BFECC semi-Lagrangian advection.
//
// Every pass is HBM-bound (sparse stencil / gather).  Reductions are deterministic:
// each block stores one partial (wave64 tree, then waves in order), and a one-workgroup
// fold kernel combines the partials in block order and updates the solver scalars, so
// the CG loop's control stays on the device and the host can enqueue iterations ahead.
// Elementwise epilogues are fused into the producing pass (p = z + beta p into Ap,
// mean subtraction into the next consumer, limiter + relaxation + mean partial).
//
// T is the state storage type (double, or float in reference-precision mode); per-cell
// arithmetic is double as in the reference (src/fluid.c).
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256
#define FIN_BS 1024
#ifndef TILED_AP
#define TILED_AP 1
#endif

template <typename T>
__device__ __forceinline__ double applyK_at(const ggbk_grid &g, const T *x, int i)
{
	// (K x)_i = sum_k w_ik (x_i - x_nb(k)) in stored neighbour order (fluid.c:277-286)
	double xi = (double) x[i], s = 0;
	for_stencil(g, i, [&](int j, float w) { s += (double) w * (xi - (double) x[j]); });
	return s;
}

__device__ __forceinline__ void store_partial(ggbk_red red, int slot, double v)
{
	double b = block_sum<BS>(v);
	if (threadIdx.x == 0)
		red.part[(size_t) slot * red.stride + xblock()] = b;
}

__device__ __forceinline__ void store_partial_max(ggbk_red red, int slot, double v)
{
	double b = block_max<BS>(v);
	if (threadIdx.x == 0)
		red.part[(size_t) slot * red.stride + xblock()] = b;
}

// ---------------- fold kernel ----------------

// Fold nb partials: thread t sums p[t], p[t + FIN_BS], ... in order, then the block
// tree; fixed order for a given nb.  Loads are issued eight at a time (the last batch
// from clamped indices, masked), so a fold of up to 8 FIN_BS partials costs one round
// trip to L2.
#define FOLD_BATCH 8
template <bool MAX>
__device__ __forceinline__ double fold(const double *p, int nb)
{
	double acc = 0;
	int i = threadIdx.x;
	for (; i < nb; i += FOLD_BATCH * FIN_BS) {
		double v[FOLD_BATCH];
#pragma unroll
		for (int q = 0; q < FOLD_BATCH; q++)
			v[q] = p[min(i + q * FIN_BS, nb - 1)];
#pragma unroll
		for (int q = 0; q < FOLD_BATCH; q++)
			if (i + q * FIN_BS < nb)
				acc = MAX ? fmax(acc, v[q]) : acc + v[q];
	}
	return MAX ? block_max<FIN_BS>(acc) : block_sum<FIN_BS>(acc);
}

__device__ __forceinline__ double fold_sum(const double *p, int nb) { return fold<false>(p, nb); }
__device__ __forceinline__ double fold_max(const double *p, int nb) { return fold<true>(p, nb); }

// A solve's start from its bb and rr (reference fluid.c:306-314): the tolerance, and
// whether there is anything to iterate (bb == 0 is status 2: psi is zeroed after the solve)
__device__ __forceinline__ void start_fold(ggbk_scalars *s, double bb, double rr, double tol, int maxit)
{
	s->bb = bb;
	s->rr = rr;
	s->it = 0;
	s->maxit = maxit;
	s->rz = 0;
	s->stop = 0;
	if (bb == 0.0) {
		s->status = 2;
		s->tol2 = 0;
		s->active = 0;
	} else {
		s->status = 0;
		s->tol2 = tol * tol * bb;
		s->active = (rr > s->tol2 && maxit > 0) ? 1 : 0;
	}
	s->vc = s->active;
}

// The mixed-precision PCG's fold (ggbk_fold_cheb): workgroup q folds the partials of sum q
// (fold_sum's fixed order) into slot G0 + q; the last workgroup to finish, told by a ticket
// taken after a device-scope fence, reads the nine sums and updates the scalars.  The ticket
// counter is reset by that workgroup; nobody waits.  With nb0 > 0 workgroups 9 and 10 fold
// the start's bb and rr partials, and the last workgroup starts the solve first.
__global__ __launch_bounds__(FIN_BS) void k_fold_cheb(ggbk_scalars *s, ggbk_red red, int nb, int first, int nb0,
						      double tol, int maxit)
{
	__shared__ int last;
	const int q = blockIdx.x, nwg = gridDim.x;
	const int slot = q == 0 ? GGBK_SL_RZ : q == 8 ? GGBK_SL_RR : q == 9 ? GGBK_SL_BB : q == 10 ? GGBK_SL_RR0
											: GGBK_SL_ZAZ + q - 1;
	const size_t st = (size_t) red.stride;
	double v = fold_sum(red.part + slot * st, q >= 9 ? nb0 : nb);
	if (threadIdx.x == 0)
		red.part[(GGBK_SL_G0 + q) * st] = v;
	__threadfence();
	__syncthreads();
	if (threadIdx.x == 0)
		last = atomicAdd(&red.cnt[0], 1u) == (unsigned) (nwg - 1);
	__syncthreads();
	if (!last || threadIdx.x != 0)
		return;
	__threadfence();
	double vv[9];
#pragma unroll
	for (int k = 0; k < 9; k++)
		vv[k] = red.part[(GGBK_SL_G0 + k) * st];
	red.cnt[0] = 0;
	if (nwg == 11)
		start_fold(s, red.part[(GGBK_SL_G0 + 9) * st], red.part[(GGBK_SL_G0 + 10) * st], tol, maxit);
	ggbk_cheb_fold(s, vv, first);
}

extern "C" void ggbk_fold_cheb(ggbk_scalars *s, ggbk_red red, int nb, int first, int nb0, double tol, int maxit)
{
	GGBK_LAUNCH(k_fold_cheb, nb0 > 0 ? 11 : 9, FIN_BS, 0, s, red, nb, first, nb0, tol, maxit);
}

__global__ __launch_bounds__(FIN_BS) void k_fin(int op, ggbk_scalars *s, ggbk_red red, int nb, double tol, int maxit,
						 double total_area)
{
	// the scalars an update reads are loaded together with the partials, and the loop
	// flag is checked only after the fold, so the kernel waits for memory once
	bool guarded = op == GGBK_FIN_PAP || op == GGBK_FIN_RR || op == GGBK_FIN_RZ_FIRST || op == GGBK_FIN_RZ;
	if (op == GGBK_FIN_CHEB_END) {
		if (threadIdx.x == 0 && s->active && s->stop) {
			s->active = 0;
			s->vc = 0;
		}
		return;
	}
	int act = s->active;
	double rz = s->rz, tol2 = s->tol2;
	int it = s->it, smaxit = s->maxit;
	const double *P = red.part;
	size_t st = red.stride;
	double a = 0, b = 0;
	switch (op) {
	case GGBK_FIN_START:
		a = fold_sum(P + GGBK_SL_BB * st, nb);
		b = fold_sum(P + GGBK_SL_RR0 * st, nb);
		break;
	case GGBK_FIN_PAP:
		a = fold_sum(P + GGBK_SL_PAP * st, nb);
		break;
	case GGBK_FIN_RR:
		a = fold_sum(P + GGBK_SL_RR * st, nb);
		break;
	case GGBK_FIN_RZ_FIRST:
	case GGBK_FIN_RZ:
		a = fold_sum(P + GGBK_SL_RZ * st, nb);
		break;
	case GGBK_FIN_MEAN_F64:
	case GGBK_FIN_MEAN_F32:
		a = fold_sum(P + GGBK_SL_MEAN * st, nb);
		break;
	case GGBK_FIN_VMAX:
		a = fold_max(P + GGBK_SL_VMAX * st, nb);
		break;
	}
	if (threadIdx.x != 0 || (guarded && !act))
		return;
	switch (op) {
	case GGBK_FIN_START:
		start_fold(s, a, b, tol, maxit);
		break;
	case GGBK_FIN_PAP: {
		double pAp = a;
		s->pAp = pAp;
		if (pAp <= 0) {
			s->status = 1;
			s->active = 0;
		} else {
			s->alpha = rz / pAp;
		}
		break;
	}
	case GGBK_FIN_RR: {
		double rr = a;
		s->rr = rr;
		it = it + 1;
		s->it = it;
		if (rr <= tol2) {
			s->active = 0;
			s->status = 0;
		} else if (it >= smaxit) {
			s->active = 0;
			s->status = 3;
		}
		break;
	}
	case GGBK_FIN_RZ_FIRST:
		s->rz = a;
		break;
	case GGBK_FIN_RZ:
		s->beta = a / rz;
		s->rz = a;
		break;
	case GGBK_FIN_MEAN_F64:
	case GGBK_FIN_MEAN_F32: {
		s->sum = a;
		double m = a / total_area;
		s->mean = op == GGBK_FIN_MEAN_F32 ? (double) (float) m : m;
		break;
	}
	case GGBK_FIN_VMAX:
		s->vmax2 = a;
		break;
	}
}

// ---------------- free operators ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_curl(ggbk_grid g, const T *vel, T *out)
{
	int i = xblock() * BS + threadIdx.x;
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
	int i = xblock() * BS + threadIdx.x;
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
	int i = xblock() * BS + threadIdx.x;
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
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		y[i] = applyK_at(g, x, i);
}

// area-weighted sum partials (fluid.c:264-275)
template <typename T>
__global__ __launch_bounds__(BS) void k_mean_partial(ggbk_grid g, const T *f, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	double v = i < g.n ? (double) g.area[i] * (double) f[i] : 0.0;
	store_partial(red, GGBK_SL_MEAN, v);
}

// After a solve: psi = 0 when b was zero (status 2, as k_zero_if_bb0), then psi's mean
// partial as k_mean_partial, in one pass
template <typename T>
__global__ __launch_bounds__(BS) void k_mean_partial_bb0(ggbk_grid g, T *psi, const ggbk_scalars *s, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	bool zero = s->status == 2;
	double v = 0.0;
	if (i < g.n) {
		if (zero)
			psi[i] = (T) 0;
		else
			v = (double) g.area[i] * (double) psi[i];
	}
	store_partial(red, GGBK_SL_MEAN, v);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_sub_mean(int n, T *f, const ggbk_scalars *s)
{
	int i = xblock() * BS + threadIdx.x;
	if (i < n)
		f[i] -= (T) s->mean;
}

// ---------------- PCG (fluid.c:288-358) ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_cg_start(ggbk_grid g, T *omega, const T *psi, T *r, int sub_mean,
						  ggbk_scalars *s, ggbk_red red, int open_vc)
{
	int i = xblock() * BS + threadIdx.x;
	double bbl = 0, rrl = 0;
	if (open_vc && i == 0)
		s->vc = 1; // no block of this pass reads it
	if (i < g.n) {
		T om = omega[i];
		if (sub_mean) {
			om -= (T) s->mean;
			omega[i] = om;
		}
		double b = -(double) g.area[i] * (double) om;
		double ri = b - applyK_at(g, psi, i);
		r[i] = (T) ri;
		bbl = b * b;
		rrl = ri * ri;
	}
	store_partial(red, GGBK_SL_BB, bbl);
	store_partial(red, GGBK_SL_RR0, rrl);
}

// p_out = z (first) or z + beta p_in, then Ap = K p_out with p_out recomputed at the
// neighbours (the update of p is fused into the stencil pass; p is double-buffered).
// FIRST is a template parameter: a run-time flag around the neighbour loads makes the
// compiler wait for each of them in turn.
template <typename T, typename TZ, bool first>
__global__ __launch_bounds__(BS) void k_cg_Ap(ggbk_grid g, const TZ *z, const T *p_in, T *p_out, T *Ap,
					       const ggbk_scalars *s, ggbk_red red)
{
	int act = s->active; // checked before the stores only
	double beta = first ? 0.0 : s->beta;
	int i = xblock() * BS + threadIdx.x;
	double v = 0;
	if (i < g.n) {
		T pi = first ? (T) z[i] : (T) ((double) z[i] + beta * (double) p_in[i]);
		double sum = 0, xi = (double) pi;
		for_stencil(g, i, [&](int j, float w) {
			T pj = first ? (T) z[j] : (T) ((double) z[j] + beta * (double) p_in[j]);
			sum += (double) w * (xi - (double) pj);
		});
		if (act) {
			p_out[i] = pi;
			Ap[i] = (T) sum;
		}
		v = sum * xi;
	}
	if (act)
		store_partial(red, GGBK_SL_PAP, v);
}

// k_cg_Ap after the first iteration on face tiles: p_out = z + beta p_in is computed once
// per tile and ring cell and staged in LDS, where the stencil reads it.
template <typename T, typename TZ>
__global__ __launch_bounds__(BS) void k_cg_Ap_tile(ggbk_grid g, const TZ *z, const T *p_in, T *p_out, T *Ap,
						    const ggbk_scalars *s, ggbk_red red)
{
	__shared__ T ps[TH * TH];
	int act = s->active;
	double beta = s->beta;
	FaceTile t = face_tile(g.N);
	auto p_cell = [&](int j) { return (T) ((double) z[j] + beta * (double) p_in[j]); };
	T pi = p_cell(t.c);
	ps[(t.ty + 1) * TH + t.tx + 1] = pi;
	stage_ring<T>(t, g.N, ps, p_cell);
	__syncthreads();
	double xi = (double) pi;
	double sum = tile_stencil_sum<T>(g, t, ps, xi, p_cell);
	if (act) {
		p_out[t.c] = pi;
		Ap[t.c] = (T) sum;
	}
	if (act)
		store_partial(red, GGBK_SL_PAP, sum * xi);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_cg_update(int n, T *psi, T *r, const T *p, const T *Ap, const ggbk_scalars *s,
						   ggbk_red red)
{
	if (!s->active)
		return;
	double alpha = s->alpha;
	int i = xblock() * BS + threadIdx.x;
	double v = 0;
	if (i < n) {
		psi[i] += (T) (alpha * (double) p[i]);
		double ri = (double) r[i] - alpha * (double) Ap[i];
		r[i] = (T) ri;
		v = ri * ri;
	}
	store_partial(red, GGBK_SL_RR, v);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_jacobi_precond(ggbk_grid g, const T *r, T *z, const ggbk_scalars *s, ggbk_red red)
{
	if (!s->active)
		return;
	int i = xblock() * BS + threadIdx.x;
	double v = 0;
	if (i < g.n) {
		T zi = g.dinv[i] * r[i];
		z[i] = zi;
		v = (double) r[i] * (double) zi;
	}
	store_partial(red, GGBK_SL_RZ, v);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_zero_if_bb0(int n, T *psi, const ggbk_scalars *s)
{
	if (s->status != 2)
		return;
	int i = xblock() * BS + threadIdx.x;
	if (i < n)
		psi[i] = (T) 0;
}

// ---------------- velocity and max speed ----------------

// [psi' = psi - mean] u = p x sum_k gc_k (psi'_j - psi'_i) (fluid.c:360-376); max |u|^2 of
// the stored values in storage arithmetic (fluid.c:480-491).
template <typename T>
__global__ __launch_bounds__(BS) void k_velocity(ggbk_grid g, const T *psi_in, T *psi_out, int sub_mean, T *vel,
						  const ggbk_scalars *s, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	double m = 0;
	if (i < g.n) {
		T mean = sub_mean ? (T) s->mean : (T) 0;
		T psii = psi_in[i];
		if (sub_mean)
			psii -= mean;
		dv3 p = cell_pos(g, i);
		dv3 gr = d3(0, 0, 0);
		double pi = (double) psii;
		int nn = g.nnb[i];
		for (int k = 0; k < nn; k++) {
			int j = g.nbr[(size_t) k * g.n + i];
			T pj = psi_in[j];
			if (sub_mean)
				pj -= mean;
			dv3 c = d3(g.gc[((size_t) k * 3 + 0) * g.n + i], g.gc[((size_t) k * 3 + 1) * g.n + i],
				   g.gc[((size_t) k * 3 + 2) * g.n + i]);
			gr = vadd(gr, vscl(c, (double) pj - pi));
		}
		if (psi_out)
			psi_out[i] = psii;
		dv3 u = vcross(p, gr);
		T u0 = (T) u.x, u1 = (T) u.y, u2 = (T) u.z;
		vel[3 * (size_t) i + 0] = u0;
		vel[3 * (size_t) i + 1] = u1;
		vel[3 * (size_t) i + 2] = u2;
		T sq = u0 * u0 + u1 * u1 + u2 * u2;
		m = (double) sq;
	}
	store_partial_max(red, GGBK_SL_VMAX, m);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_max_speed2(int n, const T *vel, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	double m = 0;
	if (i < n) {
		const T *v = vel + 3 * (size_t) i;
		T sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
		m = (double) sq;
	}
	store_partial_max(red, GGBK_SL_VMAX, m);
}

// ---------------- semi-Lagrangian advection (fluid.c:397-478) ----------------

template <typename T>
__global__ __launch_bounds__(BS) void k_departures(ggbk_grid g, const T *vel, const T *omega, T *q0, T *dep_b, T *dep_f,
						    double dt, double twoOmega, int fwd)
{
	int i = xblock() * BS + threadIdx.x;
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
	// absolute vorticity q = omega + 2 Omega y (fluid.c:435-437)
	q0[i] = (T) ((double) omega[i] + twoOmega * (double) g.pos[3 * (size_t) i + 1]);
}

template <typename T>
__global__ __launch_bounds__(BS) void k_advect(ggbk_grid g, const T *src, const T *dep, T *dst)
{
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		dst[i] = (T) sample_scalar(g, src, ld3(dep + 3 * (size_t) i));
}

// q2 = A_f q1, then the corrected source q2 = q + 0.5 (q - q2) in storage arithmetic
// (fluid.c:445-448)
template <typename T>
__global__ __launch_bounds__(BS) void k_advect_bfecc(ggbk_grid g, const T *q1, const T *q, const T *dep_f, T *q2)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	T a = (T) sample_scalar(g, q1, ld3(dep_f + 3 * (size_t) i));
	T qi = q[i];
	q2[i] = qi + (T) 0.5f * (qi - a);
}

// limited final advection (fluid.c:449-464), un-Coriolis and relaxation (fluid.c:468-474),
// and the area-weighted partial of the new omega for the mean removal
template <typename T>
__global__ __launch_bounds__(BS) void k_limited_relax(ggbk_grid g, const T *q2, const T *q, const T *dep_b, const T *q1,
						       int bfecc, const T *band, T *omega, double twoOmega, double relax,
						       ggbk_red red, double *psi, const double *h1, double *h2, double c0,
						       double c1, double c2)
{
	int i = xblock() * BS + threadIdx.x;
	double part = 0;
	if (i < g.n) {
		if (psi) {
			// the solve's warm start, psi <- c0 psi + c1 h1 + c2 h2 and h2 <- the old psi (this
			// pass writes nothing it reads)
			double a = psi[i], b = h1[i], c = h2[i];
			h2[i] = a;
			psi[i] = c0 * a + c1 * b + c2 * c;
		}
		T qn;
		if (bfecc) {
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
			qn = (T) v;
		} else {
			qn = q1[i];
		}
		double w = (double) qn - twoOmega * (double) g.pos[3 * (size_t) i + 1];
		w += relax * ((double) band[i] - w);
		T om = (T) w;
		omega[i] = om;
		part = (double) g.area[i] * (double) om;
	}
	store_partial(red, GGBK_SL_MEAN, part);
}

// ---------------- launchers ----------------

#define NB(n) ggbk_blocks((n), BS)
#define DISPATCH_T(T, fn, ...)                                     \
	do {                                                       \
		if ((T) == GGBK_F64)                               \
			GGBK_LAUNCH(fn<double>, __VA_ARGS__);      \
		else                                               \
			GGBK_LAUNCH(fn<float>, __VA_ARGS__);       \
	} while (0)

extern "C" int ggbk_nblocks(int n) { return (int) NB(n); }

extern "C" void ggbk_fin(int op, ggbk_scalars *s, ggbk_red red, int nb, double tol, int maxit, double total_area)
{
	GGBK_LAUNCH(k_fin, 1, FIN_BS, 0, op, s, red, nb, tol, maxit, total_area);
}

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

extern "C" void ggbk_mean_partial_bb0(ggbk_grid g, void *psi, int T, const ggbk_scalars *s, ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_mean_partial_bb0<double>, NB(g.n), BS, 0, g, (double *) psi, s, red);
	else
		GGBK_LAUNCH(k_mean_partial_bb0<float>, NB(g.n), BS, 0, g, (float *) psi, s, red);
}

extern "C" void ggbk_mean_partial(ggbk_grid g, const void *f, int T, ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_mean_partial<double>, NB(g.n), BS, 0, g, (const double *) f, red);
	else
		GGBK_LAUNCH(k_mean_partial<float>, NB(g.n), BS, 0, g, (const float *) f, red);
}

extern "C" void ggbk_sub_mean(int n, void *f, int T, const ggbk_scalars *s)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_sub_mean<double>, NB(n), BS, 0, n, (double *) f, s);
	else
		GGBK_LAUNCH(k_sub_mean<float>, NB(n), BS, 0, n, (float *) f, s);
}

extern "C" void ggbk_remove_mean(ggbk_grid g, void *f, int T, ggbk_scalars *s, ggbk_red red)
{
	ggbk_mean_partial(g, f, T, red);
	ggbk_fin(T == GGBK_F64 ? GGBK_FIN_MEAN_F64 : GGBK_FIN_MEAN_F32, s, red, (int) NB(g.n), 0, 0, g.total_area);
	ggbk_sub_mean(g.n, f, T, s);
}

// ---------------- scratch of the entry point that allocates nothing ----------------
// ggbk_remove_mean_noalloc's reduction scratch: the scalar block, then one per-block
// partial each, for grids of up to GGBK_NOALLOC_CELLS cells (0.2 MB; the pass and its fold
// use the MEAN slot only, so the partials are addressed with slot stride 0).  Each GPU
// holds its own copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returned on that device (the symbol's name is a host-side
// stand-in, not a device address).
#define NOALLOC_BLOCKS ((GGBK_NOALLOC_CELLS + BS - 1) / BS)
#define NOALLOC_HEAD 256
static_assert(sizeof(ggbk_scalars) <= NOALLOC_HEAD, "scalar block exceeds its scratch slot");
static __device__ __align__(256) unsigned char scratch[NOALLOC_HEAD + sizeof(double) * NOALLOC_BLOCKS];
#define NOALLOC_MAX_DEVICES 64
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];

extern "C" int ggbk_noalloc_init(void)
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
		return -1;
	void *ptr = nullptr;
	if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
		return -1;
	g_scratch_dev[dev] = ptr;
	return 0;
}

extern "C" int ggbk_remove_mean_noalloc(ggbk_grid g, float *f)
{
	if (g.n > GGBK_NOALLOC_CELLS)
		return -2;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES || !g_scratch_dev[dev])
		return -1;
	unsigned char *base = (unsigned char *) g_scratch_dev[dev];
	ggbk_scalars *s = (ggbk_scalars *) base;
	ggbk_red red = { (double *) (base + NOALLOC_HEAD), 0, nullptr };
	// stride 0: the MEAN slot only, every slot at the same place
	// every partial the fold reads is written first; the scalars start cleared
	if (ggbk_zero(s, sizeof(ggbk_scalars)))
		return -1;
	ggbk_remove_mean(g, f, GGBK_F32, s, red);
	return ggbk_check("remove_mean_noalloc") ? -1 : 0;
}

extern "C" void ggbk_cg_start(ggbk_grid g, void *omega, const void *psi, void *r, int T, int sub_mean,
			      ggbk_scalars *s, ggbk_red red, int open_vc)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_cg_start<double>, NB(g.n), BS, 0, g, (double *) omega, (const double *) psi, (double *) r,
			    sub_mean, s, red, open_vc);
	else
		GGBK_LAUNCH(k_cg_start<float>, NB(g.n), BS, 0, g, (float *) omega, (const float *) psi, (float *) r,
			    sub_mean, s, red, open_vc);
}

template <typename T, typename TZ>
static void launch_cg_Ap(ggbk_grid g, const void *z, const void *p_in, void *p_out, void *Ap, int first,
			 const ggbk_scalars *s, ggbk_red red)
{
	if (first)
		GGBK_LAUNCH((k_cg_Ap<T, TZ, true>), NB(g.n), BS, 0, g, (const TZ *) z, (const T *) p_in, (T *) p_out,
			    (T *) Ap, s, red);
	else if (TILED_AP && g.N % TT == 0)
		GGBK_LAUNCH((k_cg_Ap_tile<T, TZ>), NB(g.n), BS, 0, g, (const TZ *) z, (const T *) p_in, (T *) p_out,
			    (T *) Ap, s, red);
	else
		GGBK_LAUNCH((k_cg_Ap<T, TZ, false>), NB(g.n), BS, 0, g, (const TZ *) z, (const T *) p_in, (T *) p_out,
			    (T *) Ap, s, red);
}

extern "C" void ggbk_cg_Ap(ggbk_grid g, const void *z, int TZ, const void *p_in, void *p_out, void *Ap, int T,
			   int first, const ggbk_scalars *s, ggbk_red red)
{
	ggbk_probe_begin(GGBK_PROBE_CG_AP);
	if (T == GGBK_F64 && TZ == GGBK_F64)
		launch_cg_Ap<double, double>(g, z, p_in, p_out, Ap, first, s, red);
	else if (T == GGBK_F64)
		launch_cg_Ap<double, float>(g, z, p_in, p_out, Ap, first, s, red);
	else
		launch_cg_Ap<float, float>(g, z, p_in, p_out, Ap, first, s, red);
	ggbk_probe_end(GGBK_PROBE_CG_AP);
}

extern "C" void ggbk_cg_update(ggbk_grid g, void *psi, void *r, const void *p, const void *Ap, int T,
			       const ggbk_scalars *s, ggbk_red red)
{
	ggbk_probe_begin(GGBK_PROBE_CG_UPDATE);
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_cg_update<double>, NB(g.n), BS, 0, g.n, (double *) psi, (double *) r, (const double *) p,
			    (const double *) Ap, s, red);
	else
		GGBK_LAUNCH(k_cg_update<float>, NB(g.n), BS, 0, g.n, (float *) psi, (float *) r, (const float *) p,
			    (const float *) Ap, s, red);
	ggbk_probe_end(GGBK_PROBE_CG_UPDATE);
}

extern "C" void ggbk_jacobi_precond(ggbk_grid g, const void *r, void *z, int T, const ggbk_scalars *s, ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_jacobi_precond<double>, NB(g.n), BS, 0, g, (const double *) r, (double *) z, s, red);
	else
		GGBK_LAUNCH(k_jacobi_precond<float>, NB(g.n), BS, 0, g, (const float *) r, (float *) z, s, red);
}

// Extrapolation in time of the stream function, the warm start of the next solve:
// psi <- c0 psi + c1 h1 + c2 h2 (h1, h2 the two previous steps' psi), and h2 <- the old psi
// (the host then swaps h1 and h2, so h1 is always the newest)
extern "C" void ggbk_cg_zero_if_bb0(ggbk_grid g, void *psi, int T, const ggbk_scalars *s)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_zero_if_bb0<double>, NB(g.n), BS, 0, g.n, (double *) psi, s);
	else
		GGBK_LAUNCH(k_zero_if_bb0<float>, NB(g.n), BS, 0, g.n, (float *) psi, s);
}

extern "C" void ggbk_velocity(ggbk_grid g, const void *psi_in, void *psi_out, int sub_mean, void *vel, int T,
			      const ggbk_scalars *s, ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_velocity<double>, NB(g.n), BS, 0, g, (const double *) psi_in, (double *) psi_out, sub_mean,
			    (double *) vel, s, red);
	else
		GGBK_LAUNCH(k_velocity<float>, NB(g.n), BS, 0, g, (const float *) psi_in, (float *) psi_out, sub_mean,
			    (float *) vel, s, red);
}

extern "C" void ggbk_max_speed2(ggbk_grid g, const void *vel, int T, ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_max_speed2<double>, NB(g.n), BS, 0, g.n, (const double *) vel, red);
	else
		GGBK_LAUNCH(k_max_speed2<float>, NB(g.n), BS, 0, g.n, (const float *) vel, red);
}

extern "C" void ggbk_departures(ggbk_grid g, const void *vel, const void *omega, void *q0, void *dep_b, void *dep_f,
				int T, double dt, double twoOmega, int fwd)
{
	ggbk_probe_begin(GGBK_PROBE_DEPARTURES);
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_departures<double>, NB(g.n), BS, 0, g, (const double *) vel, (const double *) omega,
			    (double *) q0, (double *) dep_b, (double *) dep_f, dt, twoOmega, fwd);
	else
		GGBK_LAUNCH(k_departures<float>, NB(g.n), BS, 0, g, (const float *) vel, (const float *) omega,
			    (float *) q0, (float *) dep_b, (float *) dep_f, dt, twoOmega, fwd);
	ggbk_probe_end(GGBK_PROBE_DEPARTURES);
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

extern "C" void ggbk_advect_bfecc(ggbk_grid g, const void *q1, const void *q, const void *dep_f, void *q2, int T)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_advect_bfecc<double>, NB(g.n), BS, 0, g, (const double *) q1, (const double *) q,
			    (const double *) dep_f, (double *) q2);
	else
		GGBK_LAUNCH(k_advect_bfecc<float>, NB(g.n), BS, 0, g, (const float *) q1, (const float *) q,
			    (const float *) dep_f, (float *) q2);
}

extern "C" void ggbk_limited_relax(ggbk_grid g, const void *q2, const void *q, const void *dep_b, const void *q1,
				   int bfecc, const void *band, void *omega, int T, double twoOmega, double relax,
				   ggbk_red red)
{
	if (T == GGBK_F64)
		GGBK_LAUNCH(k_limited_relax<double>, NB(g.n), BS, 0, g, (const double *) q2, (const double *) q,
			    (const double *) dep_b, (const double *) q1, bfecc, (const double *) band, (double *) omega,
			    twoOmega, relax, red, (double *) nullptr, (const double *) nullptr, (double *) nullptr, 0.0, 0.0,
			    0.0);
	else
		GGBK_LAUNCH(k_limited_relax<float>, NB(g.n), BS, 0, g, (const float *) q2, (const float *) q,
			    (const float *) dep_b, (const float *) q1, bfecc, (const float *) band, (float *) omega,
			    twoOmega, relax, red, (double *) nullptr, (const double *) nullptr, (double *) nullptr, 0.0, 0.0,
			    0.0);
}

extern "C" void ggbk_limited_relax_extrap(ggbk_grid g, const double *q2, const double *q, const double *dep_b,
					  const double *q1, int bfecc, const double *band, double *omega, double twoOmega,
					  double relax, ggbk_red red, double *psi, const double *h1, double *h2, double c0,
					  double c1, double c2)
{
	GGBK_LAUNCH(k_limited_relax<double>, NB(g.n), BS, 0, g, q2, q, dep_b, q1, bfecc, band, omega, twoOmega, relax,
		    red, psi, h1, h2, c0, c1, c2);
}
