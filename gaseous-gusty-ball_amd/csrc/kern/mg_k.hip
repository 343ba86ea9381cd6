// mg_k.hip -- geometric multigrid V(1,1) preconditioner on the cubed-sphere level
// hierarchy (reference src/fluid.c:100-146): damped Jacobi (omega = 0.8), residual,
// restriction R = P^T as a CSR gather (fine index order, deterministic), bilinear
// prolongation, and the coarsest level's 60 sweeps in one LDS-resident workgroup.
//
// TM is the level storage type (double in FP64 mode, float in the mixed and reference
// modes); TB is the type of the level-0 right-hand side (the CG residual).  Per-cell
// arithmetic is double, as in the reference.
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define BS 256
#define MG_OMEGA 0.8
#define MG_COARSE_SWEEPS 60
#define COARSE_LDS_CELLS 3072
#define COARSE_BS 1024

template <typename TX>
__device__ __forceinline__ double kx_at(const ggbk_grid &g, const TX *x, int i)
{
	double xi = (double) x[i], s = 0;
	int nn = g.nnb[i];
	for (int k = 0; k < nn; k++) {
		int j = g.nbr[(size_t) k * g.n + i];
		s += (double) g.w[(size_t) k * g.n + i] * (xi - (double) x[j]);
	}
	return s;
}

// First sweep from x = 0: r = b, x = 0.8 dinv r (the stencil of a zero vector vanishes).
template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_first(ggbk_grid g, const TB *b, TM *x, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= g.n)
		return;
	TM ri = (TM) ((double) b[i] - 0.0);
	x[i] = (TM) 0 + (TM) (MG_OMEGA * (double) g.dinv[i] * (double) ri);
}

template <typename TM, typename TB>
__global__ __launch_bounds__(BS) void k_mg_residual(ggbk_grid g, const TB *b, const TM *x, TM *r,
						     const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		r[i] = (TM) ((double) b[i] - kx_at(g, x, i));
}

template <typename TM>
__global__ __launch_bounds__(BS) void k_mg_jacobi(ggbk_grid g, TM *x, const TM *r, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i < g.n)
		x[i] += (TM) (MG_OMEGA * (double) g.dinv[i] * (double) r[i]);
}

// coarse b = sum over the CSR row of rw * r_fine (product in storage precision,
// accumulated in double: fluid.c:130-136)
template <typename TM>
__global__ __launch_bounds__(BS) void k_mg_restrict(ggbk_xfer xf, const TM *rf, TM *bc, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int c = blockIdx.x * BS + threadIdx.x;
	if (c >= xf.nc)
		return;
	double acc = 0;
	int e = xf.rstart[c + 1];
	for (int k = xf.rstart[c]; k < e; k++)
		acc += (double) (xf.rw[k] * rf[xf.ridx[k]]);
	bc[c] = (TM) acc;
}

// x += P xc with the 4 stored taps (fluid.c:138-144; sum in storage precision)
template <typename TM>
__global__ __launch_bounds__(BS) void k_mg_prolong(ggbk_xfer xf, const TM *xc, TM *x, const ggbk_scalars *s)
{
	if (!s->active)
		return;
	int i = blockIdx.x * BS + threadIdx.x;
	if (i >= xf.nf)
		return;
	size_t nf = (size_t) xf.nf;
	TM v = xf.pw[i] * xc[xf.ptap[i]] + xf.pw[nf + i] * xc[xf.ptap[nf + i]] +
	       xf.pw[2 * nf + i] * xc[xf.ptap[2 * nf + i]] + xf.pw[3 * nf + i] * xc[xf.ptap[3 * nf + i]];
	x[i] += (TM) (1.0 * (double) v);
}

// Coarsest level: all sweeps in one workgroup with x and r in LDS.
template <typename TM, typename TB>
__global__ __launch_bounds__(COARSE_BS) void k_mg_coarse_lds(ggbk_grid g, const TB *b, TM *xout, const ggbk_scalars *s,
							     int sweeps)
{
	if (!s->active)
		return;
	__shared__ TM xs[COARSE_LDS_CELLS];
	__shared__ TM rs[COARSE_LDS_CELLS];
	int n = g.n;
	for (int i = threadIdx.x; i < n; i += COARSE_BS) {
		TM ri = (TM) ((double) b[i] - 0.0);
		rs[i] = ri;
		xs[i] = (TM) 0 + (TM) (MG_OMEGA * (double) g.dinv[i] * (double) ri);
	}
	__syncthreads();
	for (int sw = 1; sw < sweeps; sw++) {
		for (int i = threadIdx.x; i < n; i += COARSE_BS)
			rs[i] = (TM) ((double) b[i] - kx_at(g, xs, i));
		__syncthreads();
		for (int i = threadIdx.x; i < n; i += COARSE_BS)
			xs[i] += (TM) (MG_OMEGA * (double) g.dinv[i] * (double) rs[i]);
		__syncthreads();
	}
	for (int i = threadIdx.x; i < n; i += COARSE_BS)
		xout[i] = xs[i];
}

#define NB(n) ggbk_blocks((n), BS)

template <typename TM, typename TB>
static void smooth_from_zero(const ggbk_level &L, const TB *b, TM *x, TM *r, int sweeps, const ggbk_scalars *s)
{
	const ggbk_grid &g = L.g;
	if (sweeps > 1 && g.n <= COARSE_LDS_CELLS) {
		GGBK_LAUNCH((k_mg_coarse_lds<TM, TB>), 1, COARSE_BS, 0, g, b, x, s, sweeps);
		return;
	}
	GGBK_LAUNCH((k_mg_first<TM, TB>), NB(g.n), BS, 0, g, b, x, s);
	for (int sw = 1; sw < sweeps; sw++) {
		GGBK_LAUNCH((k_mg_residual<TM, TB>), NB(g.n), BS, 0, g, b, (const TM *) x, r, s);
		GGBK_LAUNCH((k_mg_jacobi<TM>), NB(g.n), BS, 0, g, x, (const TM *) r, s);
	}
}

template <typename TM, typename TB>
static void vcycle_level(const ggbk_level *lv, int nlev, int l, const TB *b, TM *x, TM *r, const ggbk_scalars *s)
{
	const ggbk_level &L = lv[l];
	const ggbk_grid &g = L.g;
	if (l == nlev - 1) {
		smooth_from_zero<TM, TB>(L, b, x, r, MG_COARSE_SWEEPS, s);
		return;
	}
	// pre-smooth (one sweep from zero), residual, restrict
	smooth_from_zero<TM, TB>(L, b, x, r, 1, s);
	GGBK_LAUNCH((k_mg_residual<TM, TB>), NB(g.n), BS, 0, g, b, (const TM *) x, r, s);
	const ggbk_level &C = lv[l + 1];
	GGBK_LAUNCH((k_mg_restrict<TM>), NB(L.xf.nc), BS, 0, L.xf, (const TM *) r, (TM *) C.b, s);
	vcycle_level<TM, TM>(lv, nlev, l + 1, (const TM *) C.b, (TM *) C.x, (TM *) C.r, s);
	// correct and post-smooth
	GGBK_LAUNCH((k_mg_prolong<TM>), NB(L.xf.nf), BS, 0, L.xf, (const TM *) C.x, x, s);
	GGBK_LAUNCH((k_mg_residual<TM, TB>), NB(g.n), BS, 0, g, b, (const TM *) x, r, s);
	GGBK_LAUNCH((k_mg_jacobi<TM>), NB(g.n), BS, 0, g, x, (const TM *) r, s);
}

extern "C" void ggbk_vcycle(const ggbk_level *levels, int nlev, const void *rin, int T, void *zout, int TM,
			    const ggbk_scalars *s)
{
	void *r0 = levels[0].r;
	if (TM == GGBK_F64)
		vcycle_level<double, double>(levels, nlev, 0, (const double *) rin, (double *) zout, (double *) r0, s);
	else if (T == GGBK_F64)
		vcycle_level<float, double>(levels, nlev, 0, (const double *) rin, (float *) zout, (float *) r0, s);
	else
		vcycle_level<float, float>(levels, nlev, 0, (const float *) rin, (float *) zout, (float *) r0, s);
}
