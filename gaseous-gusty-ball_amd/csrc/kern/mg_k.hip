// mg_k.hip -- geometric multigrid V(1,1) preconditioner on the cubed-sphere level
// hierarchy (reference src/fluid.c:100-146): damped Jacobi (omega = 0.8), one pre- and
// one post-sweep, restriction R = P^T, bilinear prolongation, 60 coarsest sweeps.
//
// Three passes per level instead of the reference's seven loops:
//   down     r1 = b - K x1 with x1 = 0.8 dinv b (the pre-sweep from x = 0) recomputed at
//            the neighbours, so x1 is never stored;
//   restrict b_coarse = R r1 as a padded-row gather in fine-index order (deterministic);
//   up       x = x1 + P x_coarse, then x_out = x + 0.8 dinv (b - K x), with x recomputed at
//            the neighbours (out of place), plus the r.z partial on level 0 for PCG; on
//            large levels x is stored by a prolongation pass and smoothed by a second.
// Below the first level with at most DENSE_MAX cells the V-cycle is a fixed linear map;
// it is precomputed on the host and applied as one dense matrix-vector product.
//
// The kernels of a speculative CG iteration after convergence must not store anything:
// they read the device flag `active` together with their data and check it only before
// storing, so the flag's load adds no latency to these short, latency-bound passes.
//
// TM is the level storage type (double in FP64 mode, float in the mixed and reference
// modes); TB is the type of the level-0 right-hand side (the CG residual).  Per-cell
// arithmetic is double, as in the reference.
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256
#define MG_OMEGA 0.8
#ifndef TILED_UPDATE
#define TILED_UPDATE 1
#endif

#define MG_COARSE_SWEEPS 60
#define COARSE_LDS_CELLS 3072
#define COARSE_BS 1024
#define DENSE_BS 256

// pre-sweep value of a cell with right-hand side bj: r = (TM) b, x1 = 0 + (TM)(0.8 dinv r)
template <typename TM, typename TB>
__device__ __forceinline__ TM x1_of(TB bj, float dinv)
{
	TM rj = (TM) (double) bj;
	return (TM) 0 + (TM) (MG_OMEGA * (double) dinv * (double) rj);
}

template <typename TM, typename TB>
__device__ __forceinline__ TM x1_at(const ggbk_grid &g, const TB *b, int j)
{
	return x1_of<TM, TB>(b[j], g.dinv[j]);
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_down(ggbk_grid g, const TB *b, TM *r1, const ggbk_scalars *s)
{
	int act = s->active; // waited for only at the store
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double xi = (double) x1_at<TM, TB>(g, b, i), sum = 0;
	for_stencil(g, i, [&](int j, float w) { sum += (double) w * (xi - (double) x1_at<TM, TB>(g, b, j)); });
	if (act)
		r1[i] = (TM) ((double) b[i] - sum);
}

// The PCG update and level-0 pre-smooth residual of k_cg_update_down on face tiles: each
// cell's x1 (from r_out = r_in - alpha Ap) is computed once per tile and ring cell.
template <typename T, typename TM>
__global__ __launch_bounds__(BS) void k_cg_update_down_tile(ggbk_grid g, T *psi, const T *r_in, T *r_out, const T *p,
							     const T *Ap, TM *r1, const ggbk_scalars *s, ggbk_red red)
{
	__shared__ TM x1s[TH * TH];
	int act = s->active;
	double alpha = s->alpha;
	int N = g.N;
	FaceTile t = face_tile(N);
	int i = t.c;
	double ri = (double) r_in[i] - alpha * (double) Ap[i];
	T rt = (T) ri;
	T ps = psi[i] + (T) (alpha * (double) p[i]);
	TM xi = x1_of<TM, T>(rt, g.dinv[i]);
	x1s[(t.ty + 1) * TH + t.tx + 1] = xi;
	auto x1_cell = [&](int j) {
		T rj = (T) ((double) r_in[j] - alpha * (double) Ap[j]);
		return x1_of<TM, T>(rj, g.dinv[j]);
	};
	stage_ring<TM>(t, N, x1s, x1_cell);
	__syncthreads();
	double sum = tile_stencil_sum<TM>(g, t, x1s, (double) xi, x1_cell);
	if (act) {
		psi[i] = ps;
		r_out[i] = rt;
		r1[i] = (TM) ((double) rt - sum);
	}
	double v = ri * ri;
	if (act) {
		double bsum = block_sum<BS>(v);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RR * red.stride + xblock()] = bsum;
	}
}

// k_mg_down on face tiles (x1 of each tile and ring cell staged in LDS), for level 0 where
// b is the FP64 CG residual
template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_down_tile(ggbk_grid g, const TB *b, TM *r1, const ggbk_scalars *s)
{
	__shared__ TM x1s[TH * TH];
	int act = s->active;
	FaceTile t = face_tile(g.N);
	TM xi = x1_at<TM, TB>(g, b, t.c);
	x1s[(t.ty + 1) * TH + t.tx + 1] = xi;
	auto x1_cell = [&](int j) { return x1_at<TM, TB>(g, b, j); };
	stage_ring<TM>(t, g.N, x1s, x1_cell);
	__syncthreads();
	double sum = tile_stencil_sum<TM>(g, t, x1s, (double) xi, x1_cell);
	if (act)
		r1[t.c] = (TM) ((double) b[t.c] - sum);
}

// PCG update fused with the level-0 pre-smooth residual (fluid.c:336-345 then the first
// pass of the V-cycle): psi += alpha p, r_out = r_in - alpha Ap, and r1 = r_out - K x1
// with x1 from r_out, where the neighbours' r_out is recomputed from r_in and Ap (r_out
// is a separate array, so no cell reads a value another has already replaced).  Stores
// the same values as k_cg_update followed by k_mg_down, plus the r.r partial.
template <typename T, typename TM>
__global__ __launch_bounds__(BS) void k_cg_update_down(ggbk_grid g, T *psi, const T *r_in, T *r_out, const T *p,
							const T *Ap, TM *r1, const ggbk_scalars *s, ggbk_red red)
{
	int act = s->active;
	double alpha = s->alpha;
	int i = xblock() * BS + threadIdx.x;
	double v = 0;
	if (i < g.n) {
		double ri = (double) r_in[i] - alpha * (double) Ap[i];
		T rt = (T) ri;
		T ps = psi[i] + (T) (alpha * (double) p[i]);
		double xi = (double) x1_of<TM, T>(rt, g.dinv[i]), sum = 0;
		for_stencil(g, i, [&](int j, float w) {
			T rj = (T) ((double) r_in[j] - alpha * (double) Ap[j]);
			sum += (double) w * (xi - (double) x1_of<TM, T>(rj, g.dinv[j]));
		});
		if (act) {
			psi[i] = ps;
			r_out[i] = rt;
			r1[i] = (TM) ((double) rt - sum);
		}
		v = ri * ri;
	}
	if (act) {
		double bsum = block_sum<BS>(v);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RR * red.stride + xblock()] = bsum;
	}
}

// coarse b = sum over the row of R of rw * r_fine (product in storage precision,
// accumulated in double, fine-index order: fluid.c:130-136).  Rows are padded to RP
// entries with zero weights, so every load is unconditional and independent (a bound
// known only at run time makes the compiler wait for each load in turn).
template <typename TM, int RP>
__global__ __launch_bounds__(BS) void k_mg_restrict(ggbk_xfer xf, const TM *rf, TM *bc, const ggbk_scalars *s)
{
	int act = s->active;
	int c = xblock() * BS + threadIdx.x;
	if (c >= xf.nc)
		return;
	if (xf.analytic) {
		// face-interior row: the 4x4 fine block around the coarse cell, in fine-index
		// order, with the row's weights read as four coalesced 16-byte groups
		int Nc = xf.Nc, Nf = 2 * Nc;
		int f = c / (Nc * Nc), rem = c - f * Nc * Nc;
		int J = rem / Nc, I = rem - J * Nc;
		if (I >= 1 && I <= Nc - 2 && J >= 1 && J <= Nc - 2) {
			const float4 *wq = (const float4 *) xf.rw16;
			size_t nc = (size_t) xf.nc;
			float4 wv[4];
			TM v[16];
			int i0 = (f * Nf + 2 * J - 1) * Nf + 2 * I - 1;
#pragma unroll
			for (int q = 0; q < 4; q++) {
				wv[q] = wq[q * nc + c];
				const TM *row = rf + i0 + q * Nf;
				v[4 * q] = row[0];
				v[4 * q + 1] = row[1];
				v[4 * q + 2] = row[2];
				v[4 * q + 3] = row[3];
			}
			double acc = 0;
#pragma unroll
			for (int q = 0; q < 4; q++) {
				acc += (double) (wv[q].x * v[4 * q]);
				acc += (double) (wv[q].y * v[4 * q + 1]);
				acc += (double) (wv[q].z * v[4 * q + 2]);
				acc += (double) (wv[q].w * v[4 * q + 3]);
			}
			if (act)
				bc[c] = (TM) acc;
			return;
		}
	}
	const int4 *rows = (const int4 *) xf.rr;
	size_t nc = (size_t) xf.nc;
	int idx[RP];
	float w[RP];
#pragma unroll
	for (int k = 0; k < RP; k += 2) {
		int4 v = rows[(k / 2) * nc + c];
		idx[k] = v.x;
		w[k] = __int_as_float(v.y);
		idx[k + 1] = v.z;
		w[k + 1] = __int_as_float(v.w);
	}
	double acc = 0;
#pragma unroll
	for (int k = 0; k < RP; k++)
		acc += (double) (w[k] * rf[idx[k]]);
	if (act)
		bc[c] = (TM) acc;
}

// x after prolongation at cell j: x1 + (TM)(1.0 * sum_k pw_k xc[tap_k]) with the tap
// products and sum in storage precision (fluid.c:138-144)
template <typename TM, typename TB, bool ANALYTIC = false>
__device__ __forceinline__ TM xcorr_at(const ggbk_grid &g, const ggbk_xfer &xf, const TB *b, const TM *xc, int j)
{
	TM v;
	int Nf = 2 * xf.Nc, f = 0, x = 0, y = 0;
	if (ANALYTIC) {
		f = j / (Nf * Nf);
		int rem = j - f * Nf * Nf;
		y = rem / Nf;
		x = rem - y * Nf;
	}
	if (ANALYTIC && xf.analytic && x >= 1 && x <= Nf - 2 && y >= 1 && y <= Nf - 2) {
		// face-interior fine cell: taps at (x - 1) >> 1, (y - 1) >> 1 and its right, up
		// and up-right neighbours; weights from the table
		int Nc = xf.Nc;
		int t = (f * Nc + ((y - 1) >> 1)) * Nc + ((x - 1) >> 1);
		float4 w = ((const float4 *) xf.pw)[j];
		v = w.x * xc[t] + w.y * xc[t + 1] + w.z * xc[t + Nc] + w.w * xc[t + Nc + 1];
	} else {
		const int4 *q = (const int4 *) (xf.pt + 8 * (size_t) j);
		int4 t01 = q[0], t23 = q[1];
		v = __int_as_float(t01.y) * xc[t01.x] + __int_as_float(t01.w) * xc[t01.z] +
		    __int_as_float(t23.y) * xc[t23.x] + __int_as_float(t23.w) * xc[t23.z];
	}
	return x1_at<TM, TB>(g, b, j) + (TM) (1.0 * (double) v);
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_up(ggbk_grid g, ggbk_xfer xf, const TB *b, const TM *xc, TM *xout,
					       const ggbk_scalars *s, ggbk_red red)
{
	int act = s->active;
	int i = xblock() * BS + threadIdx.x;
	double dot = 0;
	if (i < g.n) {
		TM xi = xcorr_at<TM, TB>(g, xf, b, xc, i);
		double sum = 0;
		for_stencil(g, i, [&](int j, float w) {
			sum += (double) w * ((double) xi - (double) xcorr_at<TM, TB>(g, xf, b, xc, j));
		});
		TM ri = (TM) ((double) b[i] - sum);
		TM xo = xi + (TM) (MG_OMEGA * (double) g.dinv[i] * (double) ri);
		if (act)
			xout[i] = xo;
		dot = (double) b[i] * (double) xo;
	}
	if (red.part && act) {
		double bsum = block_sum<BS>(dot);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RZ * red.stride + xblock()] = bsum;
	}
}

// Two-pass form of k_mg_up for large levels, where recomputing the prolongation at every
// neighbour costs more load instructions than storing it once:
//   prolong  xp = x1 + P x_coarse (into the level's residual array, free after restrict)
//   smooth   x_out = xp + 0.8 dinv (b - K xp), plus the r.z partial on level 0.
// Both store the same TM values as k_mg_up, so the results are identical.
template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_prolong(ggbk_grid g, ggbk_xfer xf, const TB *b, const TM *xc, TM *xp,
						    const ggbk_scalars *s)
{
	int act = s->active;
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n) {
		TM v = xcorr_at<TM, TB, true>(g, xf, b, xc, i);
		if (act)
			xp[i] = v;
	}
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_smooth(ggbk_grid g, const TB *b, const TM *xp, TM *xout,
						   const ggbk_scalars *s, ggbk_red red)
{
	int act = s->active;
	int i = xblock() * BS + threadIdx.x;
	double dot = 0;
	if (i < g.n) {
		TM xi = xp[i];
		double sum = 0;
		for_stencil(g, i, [&](int j, float w) { sum += (double) w * ((double) xi - (double) xp[j]); });
		double bi = (double) b[i];
		TM ri = (TM) (bi - sum);
		TM xo = xi + (TM) (MG_OMEGA * (double) g.dinv[i] * (double) ri);
		if (act)
			xout[i] = xo;
		dot = bi * (double) xo;
	}
	if (red.part && act) {
		double bsum = block_sum<BS>(dot);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RZ * red.stride + xblock()] = bsum;
	}
}
#define TWO_PASS_MIN_CELLS 262144

// x = M b, one workgroup per row: thread t sums columns t, t + DENSE_BS, ... (loads issued
// DENSE_UNROLL at a time), then the workgroup tree; a fixed order for a given n.  A row per
// workgroup keeps enough rows in flight to cover the load latency of this small pass.
#define DENSE_UNROLL 8
template <typename TM, typename TD, typename TB, bool VC = false>
__global__ __launch_bounds__(DENSE_BS) void k_mg_dense(const TD *M, int n, const TB *b, TM *x, const ggbk_scalars *s)
{
	int act = VC ? s->vc : s->active;
	if (VC && !act)
		return;
	int row = blockIdx.x;
	const TD *m = M + (size_t) row * n;
	double acc = 0;
	int c = threadIdx.x;
	for (; c + (DENSE_UNROLL - 1) * DENSE_BS < n; c += DENSE_UNROLL * DENSE_BS) {
		double mv[DENSE_UNROLL], bv[DENSE_UNROLL];
#pragma unroll
		for (int u = 0; u < DENSE_UNROLL; u++) {
			mv[u] = (double) m[c + u * DENSE_BS];
			bv[u] = (double) b[c + u * DENSE_BS];
		}
#pragma unroll
		for (int u = 0; u < DENSE_UNROLL; u++)
			acc += mv[u] * bv[u];
	}
	for (; c < n; c += DENSE_BS)
		acc += (double) m[c] * (double) b[c];
	acc = block_sum<DENSE_BS>(acc);
	if (threadIdx.x == 0 && act)
		x[row] = (TM) acc;
}

// Fallback for a large coarsest level: sweeps as separate passes.
template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_first(ggbk_grid g, const TB *b, TM *x, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		x[i] = x1_at<TM, TB>(g, b, i);
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_residual(ggbk_grid g, const TB *b, const TM *x, TM *r,
						     const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = xblock() * BS + threadIdx.x;
	if (i >= g.n)
		return;
	double xi = (double) x[i], sum = 0;
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++) {
		int j = g.nbr[(size_t) k * g.n + i];
		sum += (double) g.w[(size_t) k * g.n + i] * (xi - (double) x[j]);
	}
	r[i] = (TM) ((double) b[i] - sum);
}

template <typename TM>
__global__ __launch_bounds__(BS) void k_mg_jacobi(ggbk_grid g, TM *x, const TM *r, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = xblock() * BS + threadIdx.x;
	if (i < g.n)
		x[i] += (TM) (MG_OMEGA * (double) g.dinv[i] * (double) r[i]);
}

// Coarsest level with all sweeps in one workgroup, x and r in LDS.
template <typename TM, typename TB>
__global__ __launch_bounds__(COARSE_BS) void k_mg_coarse_lds(ggbk_grid g, const TB *b, TM *xout, const ggbk_scalars *s,
							     int sweeps)
{
	if (!s->active)
		return;
	__shared__ TM xs[COARSE_LDS_CELLS];
	__shared__ TM rs[COARSE_LDS_CELLS];
	int n = g.n;
	for (int i = threadIdx.x; i < n; i += COARSE_BS)
		xs[i] = x1_at<TM, TB>(g, b, i);
	__syncthreads();
	for (int sw = 1; sw < sweeps; sw++) {
		for (int i = threadIdx.x; i < n; i += COARSE_BS) {
			double xi = (double) xs[i], sum = 0;
			int nn = g.nnb[i];
			for (int k = 0; k < nn; k++) {
				int j = g.nbr[(size_t) k * g.n + i];
				sum += (double) g.w[(size_t) k * g.n + i] * (xi - (double) xs[j]);
			}
			rs[i] = (TM) ((double) b[i] - sum);
		}
		__syncthreads();
		for (int i = threadIdx.x; i < n; i += COARSE_BS)
			xs[i] += (TM) (MG_OMEGA * (double) g.dinv[i] * (double) rs[i]);
		__syncthreads();
	}
	for (int i = threadIdx.x; i < n; i += COARSE_BS)
		xout[i] = xs[i];
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_dot_partial(int n, const TB *r, const TM *z, const ggbk_scalars *s, ggbk_red red)
{
	if (!s->active)
		return;
	int i = xblock() * BS + threadIdx.x;
	double v = i < n ? (double) r[i] * (double) z[i] : 0.0;
	double bsum = block_sum<BS>(v);
	if (threadIdx.x == 0)
		red.part[(size_t) GGBK_SL_RZ * red.stride + xblock()] = bsum;
}

#define NB(n) ggbk_blocks((n), BS)

template <typename TM, typename TB>
static void coarsest(const ggbk_level &L, const TB *b, TM *x, TM *r, const ggbk_scalars *s)
{
	const ggbk_grid &g = L.g;
	if (g.n <= COARSE_LDS_CELLS) {
		GGBK_LAUNCH((k_mg_coarse_lds<TM, TB>), 1, COARSE_BS, 0, g, b, x, s, MG_COARSE_SWEEPS);
		return;
	}
	GGBK_LAUNCH((k_mg_first<TM, TB>), NB(g.n), BS, 0, g, b, x, s);
	for (int sw = 1; sw < MG_COARSE_SWEEPS; sw++) {
		GGBK_LAUNCH((k_mg_residual<TM, TB>), NB(g.n), BS, 0, g, b, (const TM *) x, r, s);
		GGBK_LAUNCH((k_mg_jacobi<TM>), NB(g.n), BS, 0, g, x, (const TM *) r, s);
	}
}

// With u != NULL the level-0 pre-smooth pass is the fused PCG update, writing b itself
// (b = u->r_in - alpha Ap), and the r.r fold follows it.
template <typename TM, typename TB>
static void vcycle_level(const ggbk_level *lv, int nlev, int l, const TB *b, TM *x, ggbk_scalars *s, ggbk_red red,
			 const ggbk_cg_fuse *u = nullptr)
{
	const ggbk_level &L = lv[l];
	const ggbk_grid &g = L.g;
	TM *r = (TM *) L.r;
	if (l > 0 && L.M) {
		unsigned nb = (unsigned) g.n;
		if (sizeof(TM) == 8)
			GGBK_LAUNCH((k_mg_dense<TM, double, TB>), nb, DENSE_BS, 0, (const double *) L.M, g.n, b, x, s);
		else
			GGBK_LAUNCH((k_mg_dense<TM, float, TB>), nb, DENSE_BS, 0, (const float *) L.M, g.n, b, x, s);
		return;
	}
	if (l == nlev - 1) {
		coarsest<TM, TB>(L, b, x, r, s);
		if (l == 0 && red.part)
			GGBK_LAUNCH((k_dot_partial<TM, TB>), NB(g.n), BS, 0, g.n, b, (const TM *) x, s, red);
		return;
	}
	const ggbk_level &C = lv[l + 1];
	if (u) {
		ggbk_probe_begin(GGBK_PROBE_CG_UPDATE);
		if (TILED_UPDATE && g.N % TT == 0)
			GGBK_LAUNCH((k_cg_update_down_tile<TB, TM>), NB(g.n), BS, 0, g, (TB *) u->psi,
				    (const TB *) u->r_in, (TB *) b, (const TB *) u->p, (const TB *) u->Ap, r, s, red);
		else
			GGBK_LAUNCH((k_cg_update_down<TB, TM>), NB(g.n), BS, 0, g, (TB *) u->psi, (const TB *) u->r_in,
				    (TB *) b, (const TB *) u->p, (const TB *) u->Ap, r, s, red);
		ggbk_probe_end(GGBK_PROBE_CG_UPDATE);
		ggbk_fin(GGBK_FIN_RR, s, red, (int) NB(g.n), 0, 0, 0);
	} else {
		if (l == 0)
			ggbk_probe_begin(GGBK_PROBE_MG_DOWN0);
		if (TILED_UPDATE && l == 0 && g.N % TT == 0)
			GGBK_LAUNCH((k_mg_down_tile<TM, TB>), NB(g.n), BS, 0, g, b, r, s);
		else
			GGBK_LAUNCH((k_mg_down<TM, TB>), NB(g.n), BS, 0, g, b, r, s);
		if (l == 0)
			ggbk_probe_end(GGBK_PROBE_MG_DOWN0);
	}
	if (L.xf.rpad == 18)
		GGBK_LAUNCH((k_mg_restrict<TM, 18>), NB(L.xf.nc), BS, 0, L.xf, (const TM *) r, (TM *) C.b, s);
	else
		GGBK_LAUNCH((k_mg_restrict<TM, GGBK_RMAX>), NB(L.xf.nc), BS, 0, L.xf, (const TM *) r, (TM *) C.b, s);
	ggbk_red none = { nullptr, 0, nullptr };
	vcycle_level<TM, TM>(lv, nlev, l + 1, (const TM *) C.b, (TM *) C.x, s, none);
	if (l == 0)
		ggbk_probe_begin(GGBK_PROBE_MG_UP0);
	if (g.n >= TWO_PASS_MIN_CELLS) {
		GGBK_LAUNCH((k_mg_prolong<TM, TB>), NB(g.n), BS, 0, g, L.xf, b, (const TM *) C.x, r, s);
		GGBK_LAUNCH((k_mg_smooth<TM, TB>), NB(g.n), BS, 0, g, b, (const TM *) r, x, s, l == 0 ? red : none);
	} else {
		GGBK_LAUNCH((k_mg_up<TM, TB>), NB(g.n), BS, 0, g, L.xf, b, (const TM *) C.x, x, s,
			    l == 0 ? red : none);
	}
	if (l == 0)
		ggbk_probe_end(GGBK_PROBE_MG_UP0);
}

extern "C" void ggbk_vcycle(const ggbk_level *levels, int nlev, const void *rin, int T, void *zout, int TM,
			    ggbk_scalars *s, ggbk_red red)
{
	if (TM == GGBK_F64)
		vcycle_level<double, double>(levels, nlev, 0, (const double *) rin, (double *) zout, s, red);
	else if (T == GGBK_F64)
		vcycle_level<float, double>(levels, nlev, 0, (const double *) rin, (float *) zout, s, red);
	else
		vcycle_level<float, float>(levels, nlev, 0, (const float *) rin, (float *) zout, s, red);
}

extern "C" int ggbk_vcycle_update(const ggbk_level *levels, int nlev, const ggbk_cg_fuse *u, void *r_out, int T,
				  void *zout, int TM, ggbk_scalars *s, ggbk_red red)
{
	if (nlev < 2 || levels[0].M)
		return -1;
	if (TM == GGBK_F64)
		vcycle_level<double, double>(levels, nlev, 0, (const double *) r_out, (double *) zout, s, red, u);
	else if (T == GGBK_F64)
		vcycle_level<float, double>(levels, nlev, 0, (const double *) r_out, (float *) zout, s, red, u);
	else
		vcycle_level<float, float>(levels, nlev, 0, (const float *) r_out, (float *) zout, s, red, u);
	return 0;
}

// The dense bottom of the mixed-precision V-cycle, whose stores s->vc gates
extern "C" void ggbk_mg_dense_vc_f32(const float *M, int n, const float *b, float *x, const ggbk_scalars *s)
{
	GGBK_LAUNCH((k_mg_dense<float, float, float, true>), (unsigned) n, DENSE_BS, 0, M, n, b, x, s);
}
