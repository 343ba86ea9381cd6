// mgt_k.hip -- the mixed-precision preconditioner: a V-cycle whose smoothing passes are
// Chebyshev polynomials in D^-1 K, each pass fused into one launch on 16x16 face tiles.
//
// Smoothing.  D^-1 K (K the Voronoi Laplacian, D its diagonal) has its spectrum in
// [0, 2] (Gershgorin: each row of D^-1 W sums to one).  The smoother is the Chebyshev
// iteration for the interval [1/2, 2], which damps the upper three quarters of the spectrum
// (the error the coarse grid cannot represent): with sigma = 5/3, delta = 3/4,
//   d_0 = (4/5) D^-1 r_0,   d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) D^-1 r_k,
//   rho_0 = 1/sigma, rho_k = 1 / (2 sigma - rho_{k-1}),   x_{k+1} = x_k + d_k,
// so d_1 = (9/41) d_0 + (40/41) D^-1 r_1 and d_2 = (9/73) d_1 + (328/365) D^-1 r_2.  Its
// first step is the damped Jacobi step (omega = 0.8) of the FP64 preconditioner.  The
// same polynomial before and after the coarse correction keeps the V-cycle symmetric.
//
// Fusion.  A pass of degree D over a tile needs D rings of cells around it: x_1 on rings
// <= D, x_2 on rings <= D - 1, ..., and the last step on the tile itself.  A workgroup
// recomputes its rings (neighbouring workgroups compute the same values: the stencil sums
// run in a fixed slot order, so the copies agree bit for bit) and keeps the iterate in LDS
// between its barriers.  Per level:
//   down   x = D steps from 0 on b, r = b - K x; stores r (for the restriction) and x;
//   up     x = x_pre + P x_coarse, then D steps; stores x.
// Level 0's down pass also applies the PCG update (psi += alpha p, r_out = r_in - alpha Ap,
// the r.r partial) and its up pass writes z and the r.z partial.
//
// Tiles.  An interior tile's rings lie on its face and every stencil there uses the regular
// offset codes, so its ext cells form a (16 + 2D)^2 square addressed by arithmetic.  A tile
// near a face seam reads host-built tables: ext cells ring by ring (BFS over the stencil
// graph) and each stencil neighbour's slot in that list.  Storage and arithmetic are float.
#include <hip/hip_runtime.h>

#include "dev.h"
#include "ggbk.h"

#define MGT_BS 256
#define MGT_T 16
#define MGT_PER 2 /* ext cells per thread: (16 + 2 D)^2 <= 512 for D <= 3 */
#define MGT_CAP (MGT_PER * MGT_BS)

// Chebyshev coefficients for [1/2, 2] (see the file header)
__device__ __forceinline__ float cheb_cd(int k)
{
	return k == 0 ? 0.0f : k == 1 ? 9.0f / 41.0f : k == 2 ? 9.0f / 73.0f : 0.0f;
}
__device__ __forceinline__ float cheb_cr(int k)
{
	return k == 0 ? 0.8f : k == 1 ? 40.0f / 41.0f : k == 2 ? 328.0f / 365.0f : 0.0f;
}

// Ext cell e of an interior tile with D rings, ring by ring: ring 0 the tile row-major,
// then each ring's top row, bottom row, left column, right column.  Returns the position
// (ey, ex) in the (16 + 2D)^2 square and the ring.
template <int D>
__device__ __forceinline__ void square_pos(int e, int &ey, int &ex, int &ring)
{
	if (e < MGT_T * MGT_T) {
		ey = D + e / MGT_T;
		ex = D + e % MGT_T;
		ring = 0;
		return;
	}
	int m = e - MGT_T * MGT_T;
#pragma unroll
	for (int r = 1; r <= D; r++) {
		int s = MGT_T + 2 * r, cnt = 4 * (s - 1);
		if (m < cnt || r == D) {
			int o = D - r; // square coordinate of the ring's first row / column
			if (m < s) {
				ey = o;
				ex = o + m;
			} else if (m < 2 * s) {
				ey = o + s - 1;
				ex = o + m - s;
			} else if (m < 2 * s + s - 2) {
				ey = o + 1 + (m - 2 * s);
				ex = o;
			} else {
				ey = o + 1 + (m - 3 * s + 2);
				ex = o + s - 1;
			}
			ring = r;
			return;
		}
		m -= cnt;
	}
}

// Per-thread view of one ext cell.
struct ExtCell {
	int c;          // cell index (-1: none)
	int slot;       // LDS slot
	int ring;
	unsigned short nb[8];
	float w[8];
	float dinv;
};

// Sets up ext cell e of tile `tile` (ring >= D: value only, no stencil).
template <int D>
__device__ __forceinline__ void ext_cell(const ggbk_grid &g, const ggbk_tiles &tl, int tile, int kind, int e, int next,
					 const int *ring_end, ExtCell &x)
{
	x.c = -1;
	if (e >= next)
		return;
	constexpr int W = MGT_T + 2 * D;
	int c;
	if (kind < 0) {
		int N = g.N, tpe = N / MGT_T;
		int f = tile / (tpe * tpe), r = tile - f * tpe * tpe;
		int J0 = (r / tpe) * MGT_T, I0 = (r % tpe) * MGT_T;
		int ey, ex, ring;
		square_pos<D>(e, ey, ex, ring);
		c = (f * N + J0 - D + ey) * N + I0 - D + ex;
		x.slot = ey * W + ex;
		x.ring = ring;
		if (ring < D) {
			unsigned code = (unsigned) ((const int2 *) g.srec)[c].y;
#pragma unroll
			for (int k = 0; k < 8; k++) {
				int dy = (int) ((code >> (4 * k + 2)) & 3u) - 1, dx = (int) ((code >> (4 * k)) & 3u) - 1;
				x.nb[k] = (unsigned short) (x.slot + dy * W + dx);
			}
		}
	} else {
		int off = tl.goff[kind];
		c = tl.gcell[off + e];
		x.slot = e;
		int ring = 0;
#pragma unroll
		for (int r = 0; r < D; r++)
			ring += e >= ring_end[r];
		x.ring = ring;
		if (ring < D) {
			const uint4 *q = (const uint4 *) (tl.gnbr + 8 * (size_t) (off + e));
			uint4 v = *q;
			x.nb[0] = (unsigned short) (v.x & 0xffff);
			x.nb[1] = (unsigned short) (v.x >> 16);
			x.nb[2] = (unsigned short) (v.y & 0xffff);
			x.nb[3] = (unsigned short) (v.y >> 16);
			x.nb[4] = (unsigned short) (v.z & 0xffff);
			x.nb[5] = (unsigned short) (v.z >> 16);
			x.nb[6] = (unsigned short) (v.w & 0xffff);
			x.nb[7] = (unsigned short) (v.w >> 16);
		}
	}
	x.c = c;
	if (x.ring < D) {
		int cls = ((const int2 *) g.srec)[c].x & 0x7fffffff;
		const float4 *wp = (const float4 *) (g.scw + 8 * (size_t) cls);
		float4 wa = wp[0], wb = wp[1];
		x.w[0] = wa.x;
		x.w[1] = wa.y;
		x.w[2] = wa.z;
		x.w[3] = wa.w;
		x.w[4] = wb.x;
		x.w[5] = wb.y;
		x.w[6] = wb.z;
		x.w[7] = wb.w;
	}
	x.dinv = g.dinv[c];
}

// K x at an ext cell from the iterate in LDS (slot order; padded slots point at the cell
// itself with weight 0)
__device__ __forceinline__ float kx_lds(const ExtCell &x, const float *xs)
{
	float xi = xs[x.slot], s = 0.0f;
#pragma unroll
	for (int k = 0; k < 8; k++)
		s += x.w[k] * (xi - xs[x.nb[k]]);
	return s;
}

// Tile index, kind and the graph tile's ext count and ring ends, shared by the workgroup.
template <int D>
__device__ __forceinline__ int tile_setup(const ggbk_tiles &tl, int tile, int &kind, int *ring_end)
{
	kind = tl.kind[tile];
	if (kind < 0) {
		constexpr int W = MGT_T + 2 * D;
		return W * W;
	}
#pragma unroll
	for (int r = 0; r <= D; r++)
		ring_end[r] = tl.gring[kind * (D + 1) + r];
	return ring_end[D];
}

// Pre-smoothing pass (down): x = D Chebyshev steps from 0 on b; r = b - K x.
// UPD: level 0 of a PCG iteration, where b = r_in - alpha Ap is the updated CG residual
// (stored as r_out on the tile, with psi += alpha p and the r.r partial).
template <int D, typename TB, bool UPD>
__global__ __launch_bounds__(MGT_BS) void k_mgt_down(ggbk_grid g, ggbk_tiles tl, const TB *b, float *r, float *xpre,
						       double *psi, const double *r_in, double *r_out, const double *p,
						       const double *Ap, const ggbk_scalars *s, ggbk_red red)
{
	__shared__ float xa[MGT_CAP], xb[MGT_CAP];
	int act = s->active;
	double alpha = UPD ? s->alpha : 0.0;
	int tile = xblock(), kind, ring_end[D + 1];
	int next = tile_setup<D>(tl, tile, kind, ring_end);
	ExtCell xc[MGT_PER];
	float bv[MGT_PER], dv[MGT_PER];
	double rr = 0;
#pragma unroll
	for (int m = 0; m < MGT_PER; m++) {
		ext_cell<D>(g, tl, tile, kind, threadIdx.x + m * MGT_BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			double bd;
			if (UPD) {
				bd = r_in[c] - alpha * Ap[c];
				if (xc[m].ring == 0) {
					double pv = psi[c] + alpha * p[c];
					if (act) {
						psi[c] = pv;
						r_out[c] = bd;
					}
					rr += bd * bd;
				}
			} else {
				bd = (double) b[c];
			}
			bv[m] = (float) bd;
			dv[m] = cheb_cr(0) * xc[m].dinv * bv[m];
			xa[xc[m].slot] = dv[m];
		}
	}
	__syncthreads();
	float *xs = xa, *xn = xb;
#pragma unroll
	for (int k = 1; k < D; k++) {
#pragma unroll
		for (int m = 0; m < MGT_PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= D - k) {
				float res = bv[m] - kx_lds(xc[m], xs);
				dv[m] = cheb_cd(k) * dv[m] + cheb_cr(k) * xc[m].dinv * res;
				xn[xc[m].slot] = xs[xc[m].slot] + dv[m];
			}
		}
		__syncthreads();
		float *t = xs;
		xs = xn;
		xn = t;
	}
#pragma unroll
	for (int m = 0; m < MGT_PER; m++) {
		if (xc[m].c >= 0 && xc[m].ring == 0 && act) {
			r[xc[m].c] = bv[m] - kx_lds(xc[m], xs);
			xpre[xc[m].c] = xs[xc[m].slot];
		}
	}
	if (UPD && act) {
		double bsum = block_sum<MGT_BS>(rr);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RR * red.stride + tile] = bsum;
	}
}

// Prolongation sum_k pw_k xc[tap_k] at fine cell c (interior fine cells of an analytic
// transfer compute their taps; the weights are always read)
__device__ __forceinline__ float prolong_at(const ggbk_xfer &xf, const float *xc, int c)
{
	int Nf = 2 * xf.Nc;
	if (xf.analytic) {
		int f = c / (Nf * Nf), rem = c - f * Nf * Nf, y = rem / Nf, x = rem - y * Nf;
		if (x >= 1 && x <= Nf - 2 && y >= 1 && y <= Nf - 2) {
			int Nc = xf.Nc;
			int t = (f * Nc + ((y - 1) >> 1)) * Nc + ((x - 1) >> 1);
			float4 w = ((const float4 *) xf.pw)[c];
			return w.x * xc[t] + w.y * xc[t + 1] + w.z * xc[t + Nc] + w.w * xc[t + Nc + 1];
		}
	}
	const int4 *q = (const int4 *) (xf.pt + 8 * (size_t) c);
	int4 t01 = q[0], t23 = q[1];
	return __int_as_float(t01.y) * xc[t01.x] + __int_as_float(t01.w) * xc[t01.z] + __int_as_float(t23.y) * xc[t23.x] +
	       __int_as_float(t23.w) * xc[t23.z];
}

// Post-smoothing pass (up): x = x_pre + P x_coarse, then D Chebyshev steps on b.
// RZ: level 0, where the result is z and the r.z partial is stored.
template <int D, typename TB, bool RZ>
__global__ __launch_bounds__(MGT_BS) void k_mgt_up(ggbk_grid g, ggbk_tiles tl, ggbk_xfer xf, const TB *b,
						     const float *xpre, const float *xcoarse, float *xout,
						     const ggbk_scalars *s, ggbk_red red)
{
	__shared__ float xa[MGT_CAP], xb[MGT_CAP];
	int act = s->active;
	int tile = xblock(), kind, ring_end[D + 1];
	int next = tile_setup<D>(tl, tile, kind, ring_end);
	ExtCell xc[MGT_PER];
	float bv[MGT_PER], dv[MGT_PER];
	double bz[MGT_PER];
#pragma unroll
	for (int m = 0; m < MGT_PER; m++) {
		ext_cell<D>(g, tl, tile, kind, threadIdx.x + m * MGT_BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		bz[m] = 0.0;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			if (xc[m].ring < D) {
				bz[m] = (double) b[c];
				bv[m] = (float) bz[m];
			}
			xa[xc[m].slot] = xpre[c] + prolong_at(xf, xcoarse, c);
		}
	}
	__syncthreads();
	float *xs = xa, *xn = xb;
#pragma unroll
	for (int k = 0; k < D; k++) {
#pragma unroll
		for (int m = 0; m < MGT_PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= D - 1 - k) {
				float res = bv[m] - kx_lds(xc[m], xs);
				dv[m] = cheb_cd(k) * dv[m] + cheb_cr(k) * xc[m].dinv * res;
				float xv = xs[xc[m].slot] + dv[m];
				if (k < D - 1)
					xn[xc[m].slot] = xv;
				else if (act)
					xout[xc[m].c] = xv;
				if (k == D - 1)
					bz[m] *= (double) xv;
			}
		}
		if (k < D - 1) {
			__syncthreads();
			float *t = xs;
			xs = xn;
			xn = t;
		}
	}
	if (RZ && act && red.part) {
		double v = 0;
#pragma unroll
		for (int m = 0; m < MGT_PER; m++)
			if (xc[m].c >= 0 && xc[m].ring == 0)
				v += bz[m];
		double bsum = block_sum<MGT_BS>(v);
		if (threadIdx.x == 0)
			red.part[(size_t) GGBK_SL_RZ * red.stride + tile] = bsum;
	}
}

// Restriction and dense bottom (mg_k.hip)
extern "C" void ggbk_mg_restrict_f32(ggbk_xfer xf, const float *rf, float *bc, const ggbk_scalars *s);
extern "C" void ggbk_mg_dense_f32(const float *M, int n, const float *b, float *x, const ggbk_scalars *s);

template <int D>
static void down_level(const ggbk_level &L, const float *b, ggbk_scalars *s)
{
	ggbk_red none = { nullptr, 0 };
	GGBK_LAUNCH((k_mgt_down<D, float, false>), L.tl.ntiles, MGT_BS, 0, L.g, L.tl, b, (float *) L.r,
		    (float *) L.xpre, nullptr, nullptr, nullptr, nullptr, nullptr, s, none);
}

template <int D>
static void up_level(const ggbk_level &L, const float *b, const float *xc, float *x, ggbk_scalars *s)
{
	ggbk_red none = { nullptr, 0 };
	GGBK_LAUNCH((k_mgt_up<D, float, false>), L.tl.ntiles, MGT_BS, 0, L.g, L.tl, L.xf, b, (const float *) L.xpre, xc,
		    x, s, none);
}

extern "C" void ggbk_vcycle_cheb(const ggbk_level *lv, int nlev, const ggbk_cg_fuse *u, const void *r, void *r_out,
				 void *zout, ggbk_scalars *s, ggbk_red red)
{
	const ggbk_level &L0 = lv[0];
	// level 0 down (degree 2), fused with the PCG update when u is given
	if (u) {
		ggbk_probe_begin(GGBK_PROBE_CG_UPDATE);
		GGBK_LAUNCH((k_mgt_down<2, double, true>), L0.tl.ntiles, MGT_BS, 0, L0.g, L0.tl, (const double *) nullptr,
			    (float *) L0.r, (float *) L0.xpre, (double *) u->psi, (const double *) u->r_in,
			    (double *) r_out, (const double *) u->p, (const double *) u->Ap, s, red);
		ggbk_probe_end(GGBK_PROBE_CG_UPDATE);
		ggbk_fin(GGBK_FIN_RR, s, red, L0.tl.ntiles, 0, 0, 0);
	} else {
		ggbk_probe_begin(GGBK_PROBE_MG_DOWN0);
		ggbk_red none = { nullptr, 0 };
		GGBK_LAUNCH((k_mgt_down<2, double, false>), L0.tl.ntiles, MGT_BS, 0, L0.g, L0.tl, (const double *) r,
			    (float *) L0.r, (float *) L0.xpre, nullptr, nullptr, nullptr, nullptr, nullptr, s, none);
		ggbk_probe_end(GGBK_PROBE_MG_DOWN0);
	}
	const double *b0 = (const double *) (u ? r_out : r);
	// restrict and smooth down to the dense level
	int l = 0;
	for (;;) {
		const ggbk_level &C = lv[l + 1];
		ggbk_mg_restrict_f32(lv[l].xf, (const float *) lv[l].r, (float *) C.b, s);
		if (C.M) {
			ggbk_mg_dense_f32((const float *) C.M, C.g.n, (const float *) C.b, (float *) C.x, s);
			break;
		}
		l++;
		if (lv[l].tl.deg == 3)
			down_level<3>(lv[l], (const float *) lv[l].b, s);
		else
			down_level<2>(lv[l], (const float *) lv[l].b, s);
	}
	// back up
	for (; l >= 1; l--) {
		const ggbk_level &L = lv[l];
		if (L.tl.deg == 3)
			up_level<3>(L, (const float *) L.b, (const float *) lv[l + 1].x, (float *) L.x, s);
		else
			up_level<2>(L, (const float *) L.b, (const float *) lv[l + 1].x, (float *) L.x, s);
	}
	ggbk_probe_begin(GGBK_PROBE_MG_UP0);
	GGBK_LAUNCH((k_mgt_up<2, double, true>), L0.tl.ntiles, MGT_BS, 0, L0.g, L0.tl, L0.xf, b0, (const float *) L0.xpre,
		    (const float *) lv[1].x, (float *) zout, s, red);
	ggbk_probe_end(GGBK_PROBE_MG_UP0);
}
