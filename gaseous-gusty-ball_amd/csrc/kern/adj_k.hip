// adj_k.hip -- discrete adjoint of the fluid step (FP64 state), for the gradient of a
// loss of the final vorticity with respect to the initial vorticity and the forcing
// strength.  Each kernel is the transpose of one forward pass (reference src/fluid.c):
//
//   relaxation + limiter (fluid.c:449-474)  k_adj_limiter    (limiter active set re-evaluated)
//   BFECC source and forward advect          k_adj_bfecc_f
//   backward advect (fluid.c:416-422)        k_adj_advect_b
//   departure points (fluid.c:397-414)       k_adj_departures
//   velocity u = p x grad psi (:360-376)     k_adj_velocity   (gather over reverse slots)
//   Poisson solve                            k_adj_rhs + the forward MG-PCG + k_adj_combine
//
// The transpose of an interpolation sends w_k * lambda to the four taps; the derivative
// of the bilinear weights with respect to the sample point gives the departure-point
// adjoint.  Those transposes run without floating-point atomics: the scatter kernels write
// per-source entries (target, value), and ggbk_adj_transpose adds each target's entries in
// source order (below), so gradients repeat bit for bit.  Great-circle moves are differentiated in closed
// form (Rodrigues step, then normalisation).  Forward intermediates are recomputed by
// the forward kernels from the taped (omega, psi) of the step.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256

__device__ __forceinline__ void store_sum(ggbk_red red, double v)
{
	double b = block_sum<BS>(v);
	if (threadIdx.x == 0)
		red.part[(size_t) GGBK_SL_MEAN * red.stride + xblock()] = b;
}

__device__ __forceinline__ void face_basis(int f, dv3 &C, dv3 &R, dv3 &U)
{
	switch (f) {
	case 0: C = d3(0, 0, 1); R = d3(1, 0, 0); U = d3(0, 1, 0); break;
	case 1: C = d3(1, 0, 0); R = d3(0, 0, -1); U = d3(0, 1, 0); break;
	case 2: C = d3(0, 0, -1); R = d3(-1, 0, 0); U = d3(0, 1, 0); break;
	case 3: C = d3(-1, 0, 0); R = d3(0, 0, 1); U = d3(0, 1, 0); break;
	case 4: C = d3(0, 1, 0); R = d3(1, 0, 0); U = d3(0, 0, -1); break;
	default: C = d3(0, -1, 0); R = d3(1, 0, 0); U = d3(0, 0, 1); break;
	}
}

// Taps at p, their bilinear weights, and the gradients of the weights with respect to
// the 3-vector p (taps fixed): x = (atan(u) + pi/4)/h - 1/2 with u = p.R / p.C gives
// dx/dp = (R - u C) / (p.C h (1 + u^2)), likewise y.
__device__ __forceinline__ void taps_grad(const ggbk_grid &g, dv3 p, Taps &t, double wt[4], dv3 dw[4])
{
	t = taps_at(g, p);
	tap_weights(t, wt);
	int f = face_of(p);
	dv3 C, R, U;
	face_basis(f, C, R, U);
	double dc, u, v;
	gnomonic(f, p, dc, u, v);
	dv3 gx = vscl(vsub(R, vscl(C, u)), 1.0 / (dc * g.h * (1 + u * u)));
	dv3 gy = vscl(vsub(U, vscl(C, v)), 1.0 / (dc * g.h * (1 + v * v)));
	double fx = t.fx, fy = t.fy;
	dw[0] = vadd(vscl(gx, -(1 - fy)), vscl(gy, -(1 - fx)));
	dw[1] = vadd(vscl(gx, 1 - fy), vscl(gy, -fx));
	dw[2] = vadd(vscl(gx, -fy), vscl(gy, 1 - fx));
	dw[3] = vadd(vscl(gx, fy), vscl(gy, fx));
}

// lambda_v of out = gc_move(p, v, s), p fixed (vec3.h:33-42 differentiated).
__device__ __forceinline__ dv3 gc_move_adj(dv3 p, dv3 v, double s, dv3 lo)
{
	dv3 t = vtan(v, p);
	double sp = vlen(t);
	if (sp < 1e-300)
		return d3(0, 0, 0);
	double th = sp * s, sn, cs;
	sincos(th, &sn, &cs);
	dv3 r = vadd(vscl(p, cs), vscl(t, sn / sp));
	double l = vlen(r);
	dv3 out = vscl(r, 1.0 / l);
	dv3 lr = vscl(vsub(lo, vscl(out, vdot(out, lo))), 1.0 / l);
	double tl = vdot(t, lr);
	dv3 lt = vscl(lr, sn / sp);
	double lth = -sn * vdot(p, lr) + cs / sp * tl;
	double lsp = -sn / (sp * sp) * tl + s * lth;
	lt = vadd(lt, vscl(t, lsp / sp));
	return vtan(lt, p);
}

// Entry e of a scatter: its target (-1: none) and W values.
struct TrEntries {
	int *tgt;
	double *val;
};

__device__ __forceinline__ void put1(const TrEntries &t, size_t e, int target, double v)
{
	t.tgt[e] = target;
	t.val[e] = v;
}

__device__ __forceinline__ void put3(const TrEntries &t, size_t e, int target, dv3 v)
{
	t.tgt[e] = target;
	st3(t.val + 3 * e, v);
}

// Sum of x (slot MEAN partials; fold with FIN_MEAN_F64: s->sum, s->mean = sum / area).
__global__ __launch_bounds__(BS) void k_adj_sum(int n, const double *x, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	store_sum(red, i < n ? x[i] : 0.0);
}

// Relaxation toward the band (fluid.c:466-474) after the mean removal (:475), and the
// limited final advection (:449-464).  lam = dL/d omega_{n+1}; s->mean = sum(lam) / area.
// Out: four entries per cell for lq (targets [0, n)) and lq2 (targets [n, 2n)), ldb
// (store), dL/dk partial.
__global__ __launch_bounds__(BS) void k_adj_limiter(ggbk_grid g, const double *q, const double *q2,
						     const double *dep_b, const double *band, const double *lam,
						     const ggbk_scalars *s, double relax, double drelax, double c,
						     int bfecc, TrEntries tr, double *ldb, ggbk_red red)
{
	int i = xblock() * BS + threadIdx.x;
	double dk = 0;
	if (i < g.n) {
		double lpp = lam[i] - (double) g.area[i] * s->mean;
		Taps t;
		double wt[4];
		dv3 dw[4];
		taps_grad(g, ld3(dep_b + 3 * (size_t) i), t, wt, dw);
		const double *src = bfecc ? q2 : q;
		double v = 0, lo = 1e30, hi = -1e30;
		int klo = 0, khi = 0;
#pragma unroll
		for (int k = 0; k < 4; k++) {
			v += wt[k] * src[t.idx[k]];
			double o = q[t.idx[k]];
			if (o < lo) {
				lo = o;
				klo = k;
			}
			if (o > hi) {
				hi = o;
				khi = k;
			}
		}
		int state = 0;
		double qn = v;
		if (bfecc) {
			if (v < lo) {
				qn = lo;
				state = 1;
			}
			if (qn > hi) {
				qn = hi;
				state = 2;
			}
		}
		double w = qn - c * (double) g.pos[3 * (size_t) i + 1];
		dk = lpp * (band[i] - w) * drelax;
		double lqn = (1 - relax) * lpp;
		dv3 d = d3(0, 0, 0);
		size_t e = 4 * (size_t) i;
		if (state == 1 || state == 2) {
			// a clamped cell passes its adjoint to the extreme tap
			put1(tr, e, t.idx[state == 1 ? klo : khi], lqn);
			for (int k = 1; k < 4; k++)
				put1(tr, e + k, -1, 0.0);
		} else {
			int off = bfecc ? g.n : 0; // lq2 when the source was the BFECC field
#pragma unroll
			for (int k = 0; k < 4; k++) {
				put1(tr, e + k, t.idx[k] + off, wt[k] * lqn);
				d = vadd(d, vscl(dw[k], src[t.idx[k]] * lqn));
			}
		}
		st3(ldb + 3 * (size_t) i, d);
	}
	store_sum(red, dk);
}

// q2 = q + 0.5 (q - A_f q1) (fluid.c:446-448): lq += 1.5 lq2; then the transpose of
// A_f applied to -0.5 lq2, as four entries per cell for lq1, and its weight derivative
// gives ldf.
__global__ __launch_bounds__(BS) void k_adj_bfecc_f(ggbk_grid g, const double *lq2, const double *q1,
						     const double *dep_f, double *lq, TrEntries tr, double *ldf)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double l2 = lq2[i];
	lq[i] += 1.5 * l2;
	double l = -0.5 * l2;
	Taps t;
	double wt[4];
	dv3 dw[4];
	taps_grad(g, ld3(dep_f + 3 * (size_t) i), t, wt, dw);
	dv3 d = d3(0, 0, 0);
#pragma unroll
	for (int k = 0; k < 4; k++) {
		put1(tr, 4 * (size_t) i + k, t.idx[k], wt[k] * l);
		d = vadd(d, vscl(dw[k], q1[t.idx[k]] * l));
	}
	st3(ldf + 3 * (size_t) i, d);
}

// q1 = A_b q: the transpose as four entries per cell for lq, the weight derivative added
// to ldb.
__global__ __launch_bounds__(BS) void k_adj_advect_b(ggbk_grid g, const double *lq1, const double *q,
						      const double *dep_b, TrEntries tr, double *ldb)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double l = lq1[i];
	Taps t;
	double wt[4];
	dv3 dw[4];
	taps_grad(g, ld3(dep_b + 3 * (size_t) i), t, wt, dw);
	dv3 d = ld3(ldb + 3 * (size_t) i);
#pragma unroll
	for (int k = 0; k < 4; k++) {
		put1(tr, 4 * (size_t) i + k, t.idx[k], wt[k] * l);
		d = vadd(d, vscl(dw[k], q[t.idx[k]] * l));
	}
	st3(ldb + 3 * (size_t) i, d);
}

// One departure point of cell i with steps (s1, s2) (fluid.c:404-412) differentiated
// with respect to the velocity field: pm = move(p, u_i, s1); r = sum w_k(pm) u[tap_k];
// um = tan_p(tan_pm(r)); dep = move(p, um, s2).
__device__ __forceinline__ void departure_adj(const ggbk_grid &g, const double *vel, int i, double s1, double s2,
					      dv3 ldep, const TrEntries &tr, size_t e)
{
	dv3 p = cell_pos(g, i), u = ld3(vel + 3 * (size_t) i);
	dv3 pm = gc_move(p, u, s1);
	Taps t;
	double wt[4];
	dv3 dw[4];
	taps_grad(g, pm, t, wt, dw);
	dv3 uk[4];
	dv3 r = d3(0, 0, 0);
#pragma unroll
	for (int k = 0; k < 4; k++) {
		uk[k] = ld3(vel + 3 * (size_t) t.idx[k]);
		r = vadd(r, vscl(uk[k], wt[k]));
	}
	dv3 sv = vtan(r, pm);
	dv3 um = vtan(sv, p);
	dv3 lum = gc_move_adj(p, um, s2, ldep);
	dv3 lsv = vtan(lum, p);
	dv3 lr = vtan(lsv, pm);
	dv3 lpm = vsub(vscl(lsv, -vdot(pm, r)), vscl(r, vdot(pm, lsv)));
#pragma unroll
	for (int k = 0; k < 4; k++) {
		put3(tr, e + k, t.idx[k], vscl(lr, wt[k]));
		lpm = vadd(lpm, vscl(dw[k], vdot(uk[k], lr)));
	}
	put3(tr, e + 4, i, gc_move_adj(p, u, s1, lpm));
}

// Ten entries per cell for lvel: the backward and (BFECC) forward departure points' four
// taps and the cell itself.
__global__ __launch_bounds__(BS) void k_adj_departures(ggbk_grid g, const double *vel, const double *ldb,
							const double *ldf, double dt, int bfecc, TrEntries tr)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	size_t e = 10 * (size_t) i;
	departure_adj(g, vel, i, -0.5 * dt, -dt, ld3(ldb + 3 * (size_t) i), tr, e);
	if (bfecc) {
		departure_adj(g, vel, i, 0.5 * dt, dt, ld3(ldf + 3 * (size_t) i), tr, e + 5);
	} else {
		for (int k = 5; k < 10; k++)
			put3(tr, e + k, -1, d3(0, 0, 0));
	}
}

// ---------------- transposes without atomics (ggbk_adj_transpose) ----------------
// Entries per target counted (integer atomics: the counts do not depend on their order),
// an exclusive scan of the counts into starts, each entry's id placed into its target's
// bucket (in any order), then each target adds its entries in ascending id -- the
// sources' order -- to its destination.  The ids of a bucket are taken smallest first by
// repeated selection (buckets hold about 4 to 10 entries).

__global__ __launch_bounds__(BS) void k_tr_count(long long nent, const int *tgt, int *cnt)
{
	long long e = (long long) blockIdx.x * BS + threadIdx.x;
	if (e < nent && tgt[e] >= 0)
		atomicAdd(&cnt[tgt[e]], 1);
}

#define SCAN_BS 1024
// Exclusive scan of a block's 1024 values (one per thread); *total gets their sum.
__device__ __forceinline__ int block_excl_scan(int v, int *total)
{
	__shared__ int wtot[SCAN_BS / 64];
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	int x = v;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		int y = __shfl_up(x, off, 64);
		if (lane >= off)
			x += y;
	}
	if (lane == 63)
		wtot[w] = x;
	__syncthreads();
	if (w == 0) {
		int t = lane < SCAN_BS / 64 ? wtot[lane] : 0;
#pragma unroll
		for (int off = 1; off < SCAN_BS / 64; off <<= 1) {
			int y = __shfl_up(t, off, 64);
			if (lane >= off)
				t += y;
		}
		if (lane < SCAN_BS / 64)
			wtot[lane] = t;
	}
	__syncthreads();
	int r = (w ? wtot[w - 1] : 0) + x - v;
	*total = wtot[SCAN_BS / 64 - 1];
	__syncthreads();
	return r;
}

// pass 1: each block's exclusive scan into start, its total into bsum
__global__ __launch_bounds__(SCAN_BS) void k_tr_scan_local(int nt, const int *cnt, int *start, int *bsum)
{
	int i = blockIdx.x * SCAN_BS + threadIdx.x, tot;
	int v = i < nt ? cnt[i] : 0;
	int x = block_excl_scan(v, &tot);
	if (i < nt)
		start[i] = x;
	if (threadIdx.x == 0)
		bsum[blockIdx.x] = tot;
}

// pass 2 (one workgroup): exclusive scan of the nb block totals in place, in chunks of
// TR_TOP per thread; start[nt] = the grand total
#define TR_TOP 16
__global__ __launch_bounds__(SCAN_BS) void k_tr_scan_top(int nb, int nt, int *bsum, int *start)
{
	int base = threadIdx.x * TR_TOP, loc = 0, tot;
	for (int k = 0; k < TR_TOP; k++)
		loc += base + k < nb ? bsum[base + k] : 0;
	int x = block_excl_scan(loc, &tot);
	for (int k = 0; k < TR_TOP; k++)
		if (base + k < nb) {
			int v = bsum[base + k];
			bsum[base + k] = x;
			x += v;
		}
	if (threadIdx.x == 0)
		start[nt] = tot;
}

// pass 3: add each block's offset
__global__ __launch_bounds__(SCAN_BS) void k_tr_scan_add(int nt, const int *bsum, int *start)
{
	int i = blockIdx.x * SCAN_BS + threadIdx.x;
	if (i < nt)
		start[i] += bsum[blockIdx.x];
}

__global__ __launch_bounds__(BS) void k_tr_place(long long nent, const int *tgt, const int *start, int *cursor,
						  int *list)
{
	long long e = (long long) blockIdx.x * BS + threadIdx.x;
	if (e >= nent)
		return;
	int t = tgt[e];
	if (t >= 0)
		list[start[t] + atomicAdd(&cursor[t], 1)] = (int) e;
}

// dest (targets [0, split) into d0, [split, nt) into d1, W values each) += the target's
// entries in ascending id
template <int W>
__global__ __launch_bounds__(BS) void k_tr_gather(int nt, int split, const int *start, const int *list,
						   const double *val, double *d0, double *d1)
{
	int j = blockIdx.x * BS + threadIdx.x;
	if (j >= nt)
		return;
	int s = start[j], m = start[j + 1] - s;
	if (m == 0)
		return;
	double acc[W];
#pragma unroll
	for (int w = 0; w < W; w++)
		acc[w] = 0.0;
	int prev = -1;
	for (int k = 0; k < m; k++) {
		int best = INT_MAX;
		for (int q = 0; q < m; q++) {
			int e = list[s + q];
			if (e > prev && e < best)
				best = e;
		}
		prev = best;
#pragma unroll
		for (int w = 0; w < W; w++)
			acc[w] += val[(size_t) best * W + w];
	}
	double *d = j < split ? d0 + (size_t) j * W : d1 + (size_t) (j - split) * W;
#pragma unroll
	for (int w = 0; w < W; w++)
		d[w] += acc[w];
}

// vel_i = p_i x sum_k gc_ik (psi_nbr(i,k) - psi_i): with c_ik = gc_ik . (lvel_i x p_i),
// lpsi_j = sum over the cells i that list j (slot rslot) of c_i,slot - sum_k c_jk.  The
// stencil is reciprocal, so the transpose is a gather; plus the sum partial of lpsi.
__global__ __launch_bounds__(BS) void k_adj_velocity(ggbk_grid g, const double *lvel, const unsigned char *rslot,
						      double *lpsi, ggbk_red red)
{
	int j = xblock() * BS + threadIdx.x;
	double acc = 0;
	if (j < g.n) {
		size_t n = (size_t) g.n;
		dv3 mj = vcross(ld3(lvel + 3 * (size_t) j), cell_pos(g, j));
		int nn = g.nnb[j];
		for (int m = 0; m < nn; m++) {
			int i = g.nbr[m * n + j];
			int k = rslot[m * n + j];
			dv3 mi = vcross(ld3(lvel + 3 * (size_t) i), cell_pos(g, i));
			dv3 gci = d3(g.gc[(k * 3 + 0) * n + i], g.gc[(k * 3 + 1) * n + i], g.gc[(k * 3 + 2) * n + i]);
			dv3 gcj = d3(g.gc[(m * 3 + 0) * n + j], g.gc[(m * 3 + 1) * n + j], g.gc[(m * 3 + 2) * n + j]);
			acc += vdot(gci, mi) - vdot(gcj, mj);
		}
		lpsi[j] = acc;
	}
	store_sum(red, acc);
}

// Right-hand side of the adjoint Poisson solve in the solver's omega form
// (b = -area * omega): rhs = lpsi - area * S / total_area - S (1 - sum(area) / total_area) / n,
// i.e. the transpose of psi's mean removal followed by the removal of the constant
// component, which K^+ ignores.  S = s->sum.
__global__ __launch_bounds__(BS) void k_adj_rhs(ggbk_grid g, const double *lpsi, const ggbk_scalars *s, double sum_area,
						 double *om)
{
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double S = s->sum;
	double a = (double) g.area[i];
	double rhs = lpsi[i] - a * (S / g.total_area) - S * (1.0 - sum_area / g.total_area) / g.n;
	om[i] = -rhs / a;
}

// lam = lq - area * z  (lambda of omega_n: direct path plus the Poisson path)
__global__ __launch_bounds__(BS) void k_adj_combine(ggbk_grid g, const double *lq, const double *z, double *lam)
{
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		lam[i] = lq[i] - (double) g.area[i] * z[i];
}

// out = lam - area * s->mean (transpose of the area-weighted mean removal)
__global__ __launch_bounds__(BS) void k_adj_project(ggbk_grid g, const double *lam, const ggbk_scalars *s, double *out)
{
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		out[i] = lam[i] - (double) g.area[i] * s->mean;
}

__global__ void k_adj_accum(const ggbk_scalars *s, double *acc)
{
	if (threadIdx.x == 0 && blockIdx.x == 0)
		acc[0] += s->sum;
}

#define NB(n) ggbk_blocks((n), BS)

extern "C" void ggbk_adj_sum(ggbk_grid g, const double *x, ggbk_red red)
{
	GGBK_LAUNCH(k_adj_sum, NB(g.n), BS, 0, g.n, x, red);
}

static TrEntries entries(const ggbk_adj_tr *t) { return TrEntries{ t->tgt, t->val }; }

extern "C" void ggbk_adj_limiter(ggbk_grid g, const double *q, const double *q2, const double *dep_b,
				 const double *band, const double *lam, const ggbk_scalars *s, double relax,
				 double drelax, double c, int bfecc, const ggbk_adj_tr *tr, double *ldb, ggbk_red red)
{
	GGBK_LAUNCH(k_adj_limiter, NB(g.n), BS, 0, g, q, q2, dep_b, band, lam, s, relax, drelax, c, bfecc,
		    entries(tr), ldb, red);
}

extern "C" void ggbk_adj_bfecc_f(ggbk_grid g, const double *lq2, const double *q1, const double *dep_f, double *lq,
				 const ggbk_adj_tr *tr, double *ldf)
{
	GGBK_LAUNCH(k_adj_bfecc_f, NB(g.n), BS, 0, g, lq2, q1, dep_f, lq, entries(tr), ldf);
}

extern "C" void ggbk_adj_advect_b(ggbk_grid g, const double *lq1, const double *q, const double *dep_b,
				  const ggbk_adj_tr *tr, double *ldb)
{
	GGBK_LAUNCH(k_adj_advect_b, NB(g.n), BS, 0, g, lq1, q, dep_b, entries(tr), ldb);
}

extern "C" void ggbk_adj_departures(ggbk_grid g, const double *vel, const double *ldb, const double *ldf, double dt,
				    int bfecc, const ggbk_adj_tr *tr)
{
	GGBK_LAUNCH(k_adj_departures, NB(g.n), BS, 0, g, vel, ldb, ldf, dt, bfecc, entries(tr));
}

extern "C" int ggbk_adj_transpose(const ggbk_adj_tr *t, long long nent, int W, int nt, int split, double *d0,
				  double *d1)
{
	int nb = (nt + SCAN_BS - 1) / SCAN_BS;
	if (nt < 1 || nt > t->cap_targets || nent > t->cap_entries || nb > SCAN_BS * TR_TOP || (W != 1 && W != 3))
		return -1;
	if (ggbk_zero(t->cnt, sizeof(int) * (size_t) nt) || ggbk_zero(t->cursor, sizeof(int) * (size_t) nt))
		return -1;
	unsigned eb = (unsigned) ((nent + BS - 1) / BS);
	GGBK_LAUNCH(k_tr_count, eb, BS, 0, nent, (const int *) t->tgt, t->cnt);
	GGBK_LAUNCH(k_tr_scan_local, (unsigned) nb, SCAN_BS, 0, nt, (const int *) t->cnt, t->start, t->bsum);
	GGBK_LAUNCH(k_tr_scan_top, 1, SCAN_BS, 0, nb, nt, t->bsum, t->start);
	GGBK_LAUNCH(k_tr_scan_add, (unsigned) nb, SCAN_BS, 0, nt, (const int *) t->bsum, t->start);
	GGBK_LAUNCH(k_tr_place, eb, BS, 0, nent, (const int *) t->tgt, (const int *) t->start, t->cursor, t->list);
	if (W == 1)
		GGBK_LAUNCH((k_tr_gather<1>), NB(nt), BS, 0, nt, split, (const int *) t->start, (const int *) t->list,
			    (const double *) t->val, d0, d1);
	else
		GGBK_LAUNCH((k_tr_gather<3>), NB(nt), BS, 0, nt, split, (const int *) t->start, (const int *) t->list,
			    (const double *) t->val, d0, d1);
	return 0;
}

extern "C" void ggbk_adj_velocity(ggbk_grid g, const double *lvel, const unsigned char *rslot, double *lpsi,
				  ggbk_red red)
{
	GGBK_LAUNCH(k_adj_velocity, NB(g.n), BS, 0, g, lvel, rslot, lpsi, red);
}

extern "C" void ggbk_adj_rhs(ggbk_grid g, const double *lpsi, const ggbk_scalars *s, double sum_area, double *om)
{
	GGBK_LAUNCH(k_adj_rhs, NB(g.n), BS, 0, g, lpsi, s, sum_area, om);
}

extern "C" void ggbk_adj_combine(ggbk_grid g, const double *lq, const double *z, double *lam)
{
	GGBK_LAUNCH(k_adj_combine, NB(g.n), BS, 0, g, lq, z, lam);
}

extern "C" void ggbk_adj_project(ggbk_grid g, const double *lam, const ggbk_scalars *s, double *out)
{
	GGBK_LAUNCH(k_adj_project, NB(g.n), BS, 0, g, lam, s, out);
}

extern "C" void ggbk_adj_accum(const ggbk_scalars *s, double *acc)
{
	GGBK_LAUNCH(k_adj_accum, 1, 64, 0, s, acc);
}
