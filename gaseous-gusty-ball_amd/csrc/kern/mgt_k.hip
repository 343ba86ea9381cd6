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

// Sets up ext cell e of tile `tile`: an interior tile's HS-ring square, or a graph tile's
// list (rings 0..HG, from ext cell `off`); cells on rings < its tile's ring count get
// their stencil.
template <int HS, int HG>
__device__ __forceinline__ void ext_cell(const ggbk_grid &g, const ggbk_tiles &tl, int tile, int kind, int off, int e,
					 int next, const int *ring_end, ExtCell &x)
{
	x.c = -1;
	if (e >= next)
		return;
	constexpr int W = MGT_T + 2 * HS;
	int c, nst, cls = 0;
	if (kind < 0) {
		int N = g.N, tpe = N / MGT_T;
		int f = tile / (tpe * tpe), r = tile - f * tpe * tpe;
		int J0 = (r / tpe) * MGT_T, I0 = (r % tpe) * MGT_T;
		int ey, ex, ring;
		square_pos<HS>(e, ey, ex, ring);
		c = (f * N + J0 - HS + ey) * N + I0 - HS + ex;
		x.slot = ey * W + ex;
		x.ring = ring;
		nst = HS;
		if (ring < HS) {
			int2 rc = ((const int2 *) g.srec)[c];
			cls = rc.x & 0x7fffffff;
			unsigned code = (unsigned) rc.y;
#pragma unroll
			for (int k = 0; k < 8; k++) {
				int dy = (int) ((code >> (4 * k + 2)) & 3u) - 1, dx = (int) ((code >> (4 * k)) & 3u) - 1;
				x.nb[k] = (unsigned short) (x.slot + dy * W + dx);
			}
		}
	} else {
		int2 cc = ((const int2 *) tl.gcc)[off + e];
		c = cc.x;
		cls = cc.y;
		x.slot = e;
		int ring = 0;
#pragma unroll
		for (int r = 0; r < HG; r++)
			ring += e >= ring_end[r];
		x.ring = ring;
		nst = HG;
		if (ring < HG) {
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
	if (x.ring < nst) {
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

// The tile's kind and ext count, and a graph tile's first ext cell and ring ends (rings
// 0..HG).  With tl.geo the interior tiles are told by their index, without a load.
template <int HS, int HG>
__device__ __forceinline__ int tile_setup(const ggbk_grid &g, const ggbk_tiles &tl, int tile, int &kind, int &off,
					  int *ring_end)
{
	constexpr int W = MGT_T + 2 * HS;
	off = 0;
	if (tl.geo) {
		int N = g.N, tpe = N / MGT_T;
		int f = tile / (tpe * tpe), r = tile - f * tpe * tpe;
		int J0 = (r / tpe) * MGT_T, I0 = (r % tpe) * MGT_T;
		if (I0 >= HS && J0 >= HS && I0 + MGT_T - 1 + HS <= N - 1 && J0 + MGT_T - 1 + HS <= N - 1) {
			kind = -1;
			return W * W;
		}
	}
	const int4 *ti = (const int4 *) (tl.tinfo + 8 * (size_t) tile);
	int4 a = ti[0], b = ti[1];
	kind = a.x;
	if (kind < 0)
		return W * W;
	off = a.y;
	int rv[5] = { a.w, b.x, b.y, b.z, b.w };
#pragma unroll
	for (int r = 0; r <= HG; r++)
		ring_end[r] = rv[r];
	return a.z;
}

// Pre-smoothing pass with the restriction into the next level fused (down).  The
// workgroup owns an 8x8 block of coarse cells, whose rows of R cover its 16x16 fine tile
// and one ring of fine cells: x = D Chebyshev steps from 0 on b there and on D rings
// around (D + 1 rings of an interior tile's square, D around a graph tile's support), the
// fine residual r = b - K x on that support, and each coarse cell's b = sum of its row of
// R times r (float, in row order).  x is stored for the fine tile (the up pass starts from
// it).  UPD: level 0 of a PCG iteration, where b = r_in - alpha Ap is the updated CG
// residual (stored as r_out on the tile, with psi += alpha p and the r.r partial).
template <int D, typename TB, bool UPD, int PER>
__global__ __launch_bounds__(MGT_BS) void k_mgt_down(ggbk_grid g, ggbk_tiles tl, ggbk_xfer xf, const TB *b, float *bc,
						       float *xpre, double *psi, const double *r_in, double *r_out,
						       const double *p, const double *Ap, const ggbk_scalars *s,
						       ggbk_red red)
{
	constexpr int CAP = PER * MGT_BS, HS = D + 1, W = MGT_T + 2 * HS;
	__shared__ float xa[CAP], xb[CAP];
	int act = s->active;
	double alpha = UPD ? s->alpha : 0.0;
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<HS, D>(g, tl, tile, kind, off, ring_end);
	int hk = kind < 0 ? HS : D; // rings of this tile's ext set
	ExtCell xc[PER];
	float bv[PER], dv[PER];
	double rr = 0;
#pragma unroll
	for (int m = 0; m < PER; m++) {
		int e = threadIdx.x + m * MGT_BS;
		ext_cell<HS, D>(g, tl, tile, kind, off, e, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			bool own = kind < 0 ? xc[m].ring == 0 : e < MGT_T * MGT_T;
			double bd;
			if (UPD) {
				bd = r_in[c] - alpha * Ap[c];
				if (own) {
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
		for (int m = 0; m < PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= hk - k) {
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
	// fine residual on the support (into the free buffer) and x on the fine tile
#pragma unroll
	for (int m = 0; m < PER; m++) {
		if (xc[m].c >= 0 && xc[m].ring <= hk - D) {
			xn[xc[m].slot] = bv[m] - kx_lds(xc[m], xs);
			bool own = kind < 0 ? xc[m].ring == 0 : threadIdx.x + m * MGT_BS < MGT_T * MGT_T;
			if (own && act)
				xpre[xc[m].c] = xs[xc[m].slot];
		}
	}
	__syncthreads();
	if (threadIdx.x < 64) {
		int N = g.N, tpe = N / MGT_T, Nc = N / 2;
		int f = tile / (tpe * tpe), rt = tile - f * tpe * tpe;
		int J0 = (rt / tpe) * MGT_T, I0 = (rt % tpe) * MGT_T;
		int J = J0 / 2 + threadIdx.x / 8, I = I0 / 2 + threadIdx.x % 8;
		int cc = (f * Nc + J) * Nc + I;
		float acc = 0.0f;
		if (kind < 0) {
			// the 4x4 fine block x in [2I - 1, 2I + 2], y in [2J - 1, 2J + 2], row-major,
			// with the weights of the row as four 16-byte groups
			const float4 *wq = (const float4 *) xf.rw16;
			size_t nc = (size_t) xf.nc;
			int sy = 2 * J - 1 - (J0 - HS), sx = 2 * I - 1 - (I0 - HS);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				float4 wv = wq[q * nc + cc];
				const float *row = xn + (sy + q) * W + sx;
				acc += wv.x * row[0];
				acc += wv.y * row[1];
				acc += wv.z * row[2];
				acc += wv.w * row[3];
			}
		} else {
			size_t o = ((size_t) kind * 64 + threadIdx.x) * GGBK_TROW;
#pragma unroll 4
			for (int k = 0; k < GGBK_TROW; k++)
				acc += tl.grw[o + k] * xn[tl.grow[o + k]];
		}
		if (act)
			bc[cc] = acc;
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
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<D, D>(g, tl, tile, kind, off, ring_end);
	ExtCell xc[MGT_PER];
	float bv[MGT_PER], dv[MGT_PER];
	double bz[MGT_PER];
#pragma unroll
	for (int m = 0; m < MGT_PER; m++) {
		ext_cell<D, D>(g, tl, tile, kind, off, threadIdx.x + m * MGT_BS, next, ring_end, xc[m]);
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

// Dense bottom (mg_k.hip)
extern "C" void ggbk_mg_dense_f32(const float *M, int n, const float *b, float *x, const ggbk_scalars *s);

// down pass of level L into level C's b
template <int D, int PER>
static void down_level(const ggbk_level &L, const ggbk_level &C, const float *b, ggbk_scalars *s)
{
	ggbk_red none = { nullptr, 0 };
	GGBK_LAUNCH((k_mgt_down<D, float, false, PER>), L.td.ntiles, MGT_BS, 0, L.g, L.td, L.xf, b, (float *) C.b,
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
	// level 0 down (degree 2) into level 1, fused with the PCG update when u is given
	if (u) {
		ggbk_probe_begin(GGBK_PROBE_CG_UPDATE);
		GGBK_LAUNCH((k_mgt_down<2, double, true, 2>), L0.td.ntiles, MGT_BS, 0, L0.g, L0.td, L0.xf,
			    (const double *) nullptr, (float *) lv[1].b, (float *) L0.xpre, (double *) u->psi,
			    (const double *) u->r_in, (double *) r_out, (const double *) u->p, (const double *) u->Ap, s,
			    red);
		ggbk_probe_end(GGBK_PROBE_CG_UPDATE);
		ggbk_fin(GGBK_FIN_RR, s, red, L0.td.ntiles, 0, 0, 0);
	} else {
		ggbk_probe_begin(GGBK_PROBE_MG_DOWN0);
		ggbk_red none = { nullptr, 0 };
		GGBK_LAUNCH((k_mgt_down<2, double, false, 2>), L0.td.ntiles, MGT_BS, 0, L0.g, L0.td, L0.xf, (const double *) r,
			    (float *) lv[1].b, (float *) L0.xpre, nullptr, nullptr, nullptr, nullptr, nullptr, s, none);
		ggbk_probe_end(GGBK_PROBE_MG_DOWN0);
	}
	const double *b0 = (const double *) (u ? r_out : r);
	// smooth and restrict down to the dense level
	int l = 1;
	while (!lv[l].M) {
		down_level<3, 3>(lv[l], lv[l + 1], (const float *) lv[l].b, s);
		l++;
	}
	ggbk_mg_dense_f32((const float *) lv[l].M, lv[l].g.n, (const float *) lv[l].b, (float *) lv[l].x, s);
	// back up
	for (l = l - 1; l >= 1; l--) {
		const ggbk_level &L = lv[l];
		up_level<3>(L, (const float *) L.b, (const float *) lv[l + 1].x, (float *) L.x, s);
	}
	ggbk_probe_begin(GGBK_PROBE_MG_UP0);
	GGBK_LAUNCH((k_mgt_up<2, double, true>), L0.tl.ntiles, MGT_BS, 0, L0.g, L0.tl, L0.xf, b0, (const float *) L0.xpre,
		    (const float *) lv[1].x, (float *) zout, s, red);
	ggbk_probe_end(GGBK_PROBE_MG_UP0);
}
