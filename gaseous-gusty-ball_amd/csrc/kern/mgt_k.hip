// mgt_k.hip -- the mixed-precision preconditioner: a V-cycle whose smoothing passes are
// Chebyshev polynomials in D^-1 K, each pass fused into one launch on 16x16 face tiles.
//
// Smoothing.  D^-1 K (K the Voronoi Laplacian, D its diagonal) has its spectrum in
// [0, 2] (Gershgorin: the weights are non-negative and each row of D^-1 W sums to one).
// The Galerkin levels below have some negative weights; their spectra, measured, stay
// below 1.75 (tests/test_mixed_precond.py checks below 2).  The smoother is the Chebyshev
// iteration for the interval [1/2, 2], which damps the upper three quarters of the spectrum
// (the error the coarse grid cannot represent): with sigma = 5/3, delta = 3/4,
//   d_0 = (4/5) D^-1 r_0,   d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) D^-1 r_k,
//   rho_0 = 1/sigma, rho_k = 1 / (2 sigma - rho_{k-1}),   x_{k+1} = x_k + d_k,
// so d_1 = (9/41) d_0 + (40/41) D^-1 r_1 and d_2 = (9/73) d_1 + (328/365) D^-1 r_2.  Its
// first step is the damped Jacobi step (omega = 0.8) of the FP64 preconditioner.  The
// same polynomial before and after the coarse correction keeps the V-cycle symmetric.
// Level 0 (the Voronoi operator) smooths with degree 4; the coarse levels carry the
// Galerkin operators R K P (5x5 stencils, half-precision weights; host csrc/host/galerkin.c)
// and smooth with degree 2.  With those the PCG at N = 512 and tol 1e-8 needs 3 iterations
// per step where the rediscretised levels with degrees 2 / 3 needed 6.
//
// Fusion.  A pass of degree D over a tile needs D rings of cells around it: x_1 on rings
// <= D, x_2 on rings <= D - 1, ..., and the last step on the tile itself.  A workgroup
// recomputes its rings (neighbouring workgroups compute the same values: the stencil sums
// run in a fixed slot order, so the copies agree bit for bit) and keeps the iterate in LDS
// between its barriers.  Per level:
//   down   x = D steps from 0 on b, r = b - K x; stores r (for the restriction) and x;
//   up     x = x_pre + P x_coarse, then D steps; stores x.
// In the PCG, level 0's down pass also applies the update (p = z + beta p, Ap = Az + beta Ap,
// psi += alpha p, r_out = r_in - alpha Ap) and its up pass, one ring wider, also forms A z
// in double from the float z and the partials of the fold that yields the next alpha, so
// the solve needs no separate A p pass.  The passes after the last update are skipped
// (s->vc).
//
// Tiles.  An interior tile's rings lie on its face and every stencil there uses the regular
// offset codes, so its ext cells form a (16 + 2D)^2 square addressed by arithmetic.  A tile
// near a face seam reads host-built tables: ext cells ring by ring (BFS over the stencil
// graph) and each stencil neighbour's slot in that list.  Storage and arithmetic are float.
#include <hip/hip_runtime.h>
#include <string.h>

#include "dev.h"
#include "ggbk.h"

#define MGT_BS 256
#define MGT_T 16
#define MGT_PER 2 /* ext cells per thread: (16 + 2 D)^2 <= 512 for D <= 3 */
#define MGT_CAP (MGT_PER * MGT_BS)

// Chebyshev coefficients for [1/2, 2] (see the file header); rho_3 = 1095/3281 gives the
// fourth step d_3 = (369/3281) d_2 + (2920/3281) D^-1 r_3
__device__ __forceinline__ float cheb_cd(int k)
{
	return k == 0 ? 0.0f : k == 1 ? 9.0f / 41.0f : k == 2 ? 9.0f / 73.0f : k == 3 ? 369.0f / 3281.0f : 0.0f;
}
__device__ __forceinline__ float cheb_cr(int k)
{
	return k == 0 ? 0.8f : k == 1 ? 40.0f / 41.0f : k == 2 ? 328.0f / 365.0f : k == 3 ? 2920.0f / 3281.0f : 0.0f;
}

// Ext cell e of an interior tile with D rings: the (16 + 2D)^2 square row-major, so that a
// wave's lanes read runs of consecutive cells (listing the rings one after the other made
// the lanes of a ring's columns read cells a grid row apart).  Returns the position
// (ey, ex) in the square and the ring (the Chebyshev distance to the tile).
template <int D>
__device__ __forceinline__ void square_pos(int e, int &ey, int &ex, int &ring)
{
	constexpr int W = MGT_T + 2 * D;
	ey = e / W;
	ex = e - ey * W;
	int dy = max(D - ey, ey - (D + MGT_T - 1)), dx = max(D - ex, ex - (D + MGT_T - 1));
	ring = max(max(dy, dx), 0);
}

// Per-thread view of one ext cell.
struct ExtCell {
	int c;          // cell index (-1: none)
	int slot;       // LDS slot
	int ring;
	unsigned nb[4];  // stencil neighbours' LDS slots, two 16-bit slots per word
	float w[8];
	float dinv;
};

// the LDS slot of stencil neighbour k (two 16-bit slots per word)
__device__ __forceinline__ int slot_at(const unsigned *p, int k)
{
	return (int) ((p[k >> 1] >> (16 * (k & 1))) & 0xffffu);
}

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
			for (int k = 0; k < 8; k += 2) {
				int dy0 = (int) ((code >> (4 * k + 2)) & 3u) - 1, dx0 = (int) ((code >> (4 * k)) & 3u) - 1;
				int dy1 = (int) ((code >> (4 * k + 6)) & 3u) - 1, dx1 = (int) ((code >> (4 * k + 4)) & 3u) - 1;
				x.nb[k >> 1] = (unsigned) (x.slot + dy0 * W + dx0) | ((unsigned) (x.slot + dy1 * W + dx1) << 16);
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
			x.nb[0] = v.x;
			x.nb[1] = v.y;
			x.nb[2] = v.z;
			x.nb[3] = v.w;
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
		s += x.w[k] * (xi - xs[slot_at(x.nb, k)]);
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
	const int4 *ti = (const int4 *) (tl.tinfo + GGBK_TINFO * (size_t) tile);
	int4 a = ti[0], b = ti[1];
	kind = a.x;
	if (kind < 0)
		return W * W;
	off = a.y;
	int r5 = HG >= 5 ? ti[2].x : 0;
	int rv[6] = { a.w, b.x, b.y, b.z, b.w, r5 };
#pragma unroll
	for (int r = 0; r <= HG; r++)
		ring_end[r] = rv[r];
	return a.z;
}

// The restriction of a down pass: the workgroup's 64 threads each sum one row of R (its
// 8x8 block of coarse cells) over the fine residual in LDS (rn): an interior tile's 4x4
// fine block x in [2I - 1, 2I + 2], y in [2J - 1, 2J + 2] of its (16 + 2 HS)^2 square,
// with the row's weights as four 16-byte groups (exact: the products of 1/4, 3/4, 3/4, 1/4
// along each axis, computed); a graph tile's row through its slot and weight tables.
template <int HS>
__device__ __forceinline__ void restrict_tile(const ggbk_grid &g, const ggbk_tiles &tl, const ggbk_xfer &xf, int tile,
					      int kind, const float *xn, float *bc, int act)
{
	constexpr int W = MGT_T + 2 * HS;
	if (threadIdx.x < 64) {
		int N = g.N, tpe = N / MGT_T, Nc = N / 2;
		int f = tile / (tpe * tpe), rt = tile - f * tpe * tpe;
		int J0 = (rt / tpe) * MGT_T, I0 = (rt % tpe) * MGT_T;
		int J = J0 / 2 + threadIdx.x / 8, I = I0 / 2 + threadIdx.x % 8;
		int cc = (f * Nc + J) * Nc + I;
		float acc = 0.0f;
		if (kind < 0) {
			const float4 *wq = (const float4 *) xf.rw16;
			size_t nc = (size_t) xf.nc;
			int sy = 2 * J - 1 - (J0 - HS), sx = 2 * I - 1 - (I0 - HS);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				float4 wv;
				if (xf.exact) {
					float wy = (q == 0 || q == 3) ? 0.25f : 0.75f;
					wv = make_float4(wy * 0.25f, wy * 0.75f, wy * 0.75f, wy * 0.25f);
				} else {
					wv = wq[q * nc + cc];
				}
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
}

// Pre-smoothing pass with the restriction into the next level fused (down).  The
// workgroup owns an 8x8 block of coarse cells, whose rows of R cover its 16x16 fine tile
// and one ring of fine cells: x = D Chebyshev steps from 0 on b there and on D rings
// around (D + 1 rings of an interior tile's square, D around a graph tile's support), the
// fine residual r = b - K x on that support, and each coarse cell's b = sum of its row of
// R times r (float, in row order).  x is stored for the fine tile (the up pass starts from
// it).  UPD: level 0 of a PCG iteration, where b = r - alpha Ap' is the updated CG
// residual for the new direction p' = z + beta p, Ap' = Az + beta Ap (first: z, Az); the
// tile stores r_out = b, p', Ap' and psi += alpha p' (when s->active).  The smoothing's
// stores are skipped when no V-cycle follows the update (s->vc == 0).
template <int D, typename TB, bool UPD, int PER, int BS = MGT_BS>
__global__ __launch_bounds__(BS) void k_mgt_down(ggbk_grid g, ggbk_tiles tl, ggbk_xfer xf, const TB *b, float *bc,
						       float *xpre, ggbk_cg_cheb cg, const ggbk_scalars *s)
{
	constexpr int CAP = PER * BS, HS = D + 1, W = MGT_T + 2 * HS;
	__shared__ float xa[CAP], xb[CAP];
	int upd = s->active, act = s->vc; // the update; the smoothing after it
	if (!(UPD && upd) && !act)
		return;
	double alpha = UPD ? s->alpha : 0.0, beta = UPD ? s->beta : 0.0;
	bool first = cg.first;
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<HS, D>(g, tl, tile, kind, off, ring_end);
	int hk = kind < 0 ? HS : D; // rings of this tile's ext set
	ExtCell xc[PER];
	float bv[PER], dv[PER];
#pragma unroll
	for (int m = 0; m < PER; m++) {
		int e = threadIdx.x + m * BS;
		ext_cell<HS, D>(g, tl, tile, kind, off, e, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			bool own = kind < 0 ? xc[m].ring == 0 : e < MGT_T * MGT_T;
			double bd;
			if (UPD && !act) {
				// the solve's last update (no V-cycle follows): psi alone
				if (own && upd) {
					double pv = first ? (double) cg.z[c] : (double) cg.z[c] + beta * cg.p_in[c];
					cg.psi[c] = cg.psi[c] + alpha * pv;
				}
				continue;
			}
			if (UPD) {
				// p = z + beta p, Ap = Az + beta Ap (first: p = z, Ap = Az); r -= alpha Ap
				double apv = first ? cg.Az[c] : cg.Az[c] + beta * cg.Ap_in[c];
				bd = cg.r[c] - alpha * apv;
				if (own) {
					double pv = first ? (double) cg.z[c] : (double) cg.z[c] + beta * cg.p_in[c];
					double ps = cg.psi[c] + alpha * pv;
					if (upd) {
						cg.psi[c] = ps;
						cg.p_out[c] = pv;
						cg.Ap_out[c] = apv;
						cg.r_out[c] = bd;
					}
				}
			} else {
				bd = (double) b[c];
			}
			bv[m] = (float) bd;
			dv[m] = cheb_cr(0) * xc[m].dinv * bv[m];
			xa[xc[m].slot] = dv[m];
		}
	}
	if (!act)
		return; // the update was the last: no V-cycle follows
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
			bool own = kind < 0 ? xc[m].ring == 0 : threadIdx.x + m * BS < MGT_T * MGT_T;
			if (own && act)
				xpre[xc[m].c] = xs[xc[m].slot];
		}
	}
	__syncthreads();
	restrict_tile<HS>(g, tl, xf, tile, kind, xn, bc, act);
}

// The PCG update alone (ggbk_cg_update_cheb): the first pass's update without the
// smoothing, for the iteration the host expects to be the last of the solve.
__global__ __launch_bounds__(MGT_BS) void k_cg_update_cheb(int n, ggbk_cg_cheb cg, const ggbk_scalars *s)
{
	int upd = s->active, last = s->stop;
	double alpha = s->alpha, beta = s->beta;
	int c = xblock() * MGT_BS + threadIdx.x;
	if (c >= n || !upd)
		return;
	double pv = cg.first ? (double) cg.z[c] : (double) cg.z[c] + beta * cg.p_in[c];
	cg.psi[c] = cg.psi[c] + alpha * pv;
	if (last)
		return; // the solve ends here: r, p and A p are not read again
	double apv = cg.first ? cg.Az[c] : cg.Az[c] + beta * cg.Ap_in[c];
	cg.p_out[c] = pv;
	cg.Ap_out[c] = apv;
	cg.r_out[c] = cg.r[c] - alpha * apv;
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
			float4 w;
			if (xf.exact) {
				// odd x sits a quarter cell from its first tap column: 3/4 there, 1/4 on
				// the next; even x the other way round (the same for y and rows)
				float ax = (x & 1) ? 0.75f : 0.25f, ay = (y & 1) ? 0.75f : 0.25f;
				float bx = 1.0f - ax, by = 1.0f - ay;
				w = make_float4(ax * ay, bx * ay, ax * by, bx * by);
			} else {
				w = ((const float4 *) xf.pw)[c];
			}
			return w.x * xc[t] + w.y * xc[t + 1] + w.z * xc[t + Nc] + w.w * xc[t + Nc + 1];
		}
	}
	const int4 *q = (const int4 *) (xf.pt + 8 * (size_t) c);
	int4 t01 = q[0], t23 = q[1];
	return __int_as_float(t01.y) * xc[t01.x] + __int_as_float(t01.w) * xc[t01.z] + __int_as_float(t23.y) * xc[t23.x] +
	       __int_as_float(t23.w) * xc[t23.z];
}

// Post-smoothing pass (up): x = x_pre + P x_coarse, then D Chebyshev steps on b.
template <int D>
__global__ __launch_bounds__(MGT_BS) void k_mgt_up(ggbk_grid g, ggbk_tiles tl, ggbk_xfer xf, const float *b,
						     const float *xpre, const float *xcoarse, float *xout,
						     const ggbk_scalars *s)
{
	__shared__ float xa[MGT_CAP], xb[MGT_CAP];
	int act = s->vc;
	if (!act)
		return;
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<D, D>(g, tl, tile, kind, off, ring_end);
	ExtCell xc[MGT_PER];
	float bv[MGT_PER], dv[MGT_PER];
#pragma unroll
	for (int m = 0; m < MGT_PER; m++) {
		ext_cell<D, D>(g, tl, tile, kind, off, threadIdx.x + m * MGT_BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			if (xc[m].ring < D)
				bv[m] = b[c];
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
			}
		}
		if (k < D - 1) {
			__syncthreads();
			float *t = xs;
			xs = xn;
			xn = t;
		}
	}
}

// Level 0's post-smoothing pass, one ring wider than k_mgt_up: x = x_pre + P x_coarse on
// D + 1 rings, D Chebyshev steps on b (the last on rings <= 1), z stored on the tile.  With
// Az != NULL (the PCG) also A z on the tile in double from the float z (stencil slots in
// order, as the FP64 operator sums them) and the fold's partials over the tile: r.z,
// z.Az, z.Ap, r.Az, Az.Az, Az.Ap, r.Ap, Ap.Ap and r.r (slots RZ, ZAZ .. APAP, RR), Ap the
// current direction's image (zero when NULL), r = b.
// (registers held to 7 waves per SIMD: 2 us faster than the compiler's 6 at degree 4; 8
// spill and run slower)
template <int D, int PER, int BS = MGT_BS>
__global__ __launch_bounds__(BS, 7) void k_mgt_up_cg(ggbk_grid g, ggbk_tiles tl, ggbk_xfer xf, const double *b,
							const float *xpre, const float *xcoarse, float *zout, double *Az,
							const double *Ap, const ggbk_scalars *s, ggbk_red red)
{
	constexpr int H = D + 1, CAP = PER * BS;
	__shared__ float xa[CAP], xb[CAP];
	__shared__ double wsum[9][BS / 64];
	int act = s->vc;
	if (!act)
		return;
	int tile = xblock(), kind, off, ring_end[H + 1];
	int next = tile_setup<H, H>(g, tl, tile, kind, off, ring_end);
	ExtCell xc[PER];
	float bv[PER], dv[PER];
	double rd[PER];
#pragma unroll
	for (int m = 0; m < PER; m++) {
		ext_cell<H, H>(g, tl, tile, kind, off, threadIdx.x + m * BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		rd[m] = 0.0;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			if (xc[m].ring < H) {
				rd[m] = b[c];
				bv[m] = (float) rd[m];
			}
			xa[xc[m].slot] = xpre[c] + prolong_at(xf, xcoarse, c);
		}
	}
	__syncthreads();
	float *xs = xa, *xn = xb;
#pragma unroll
	for (int k = 0; k < D; k++) {
#pragma unroll
		for (int m = 0; m < PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= H - 1 - k) {
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
	// z is in xs on rings <= 1
	double v[9];
#pragma unroll
	for (int q = 0; q < 9; q++)
		v[q] = 0.0;
#pragma unroll
	for (int m = 0; m < PER; m++) {
		if (xc[m].c >= 0 && xc[m].ring == 0) {
			int c = xc[m].c;
			float zf = xs[xc[m].slot];
			if (act)
				zout[c] = zf;
			if (Az) {
				double zi = (double) zf, az = 0.0;
#pragma unroll
				for (int k = 0; k < 8; k++)
					az += (double) xc[m].w[k] * (zi - (double) xs[slot_at(xc[m].nb, k)]);
				double ap = Ap ? Ap[c] : 0.0, r = rd[m];
				if (act)
					Az[c] = az;
				v[0] += r * zi;
				v[1] += zi * az;
				v[2] += zi * ap;
				v[3] += r * az;
				v[4] += az * az;
				v[5] += az * ap;
				v[6] += r * ap;
				v[7] += ap * ap;
				v[8] += r * r;
			}
		}
	}
	if (Az && act && red.part) {
		// nine block sums in a fixed order, one barrier
		int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
		for (int q = 0; q < 9; q++) {
			double t = wave_sum(v[q]);
			if (lane == 0)
				wsum[q][wid] = t;
		}
		__syncthreads();
		if (threadIdx.x < 9) {
			int q = threadIdx.x;
			double t = 0.0;
#pragma unroll
			for (int w = 0; w < BS / 64; w++)
				t += wsum[q][w];
			int slot = q == 0 ? GGBK_SL_RZ : q == 8 ? GGBK_SL_RR : GGBK_SL_ZAZ + q - 1;
			red.part[(size_t) slot * red.stride + tile] = t;
		}
	}
}

// ---- Galerkin coarse levels (wide stencils) ----
//
// A coarse level of the hierarchy carries K_c = R K_f P (ggbk_wide): a 5x5 stencil on the
// face interior, 24 weights per cell in half precision.  A smoothing step then needs two
// rings of cells around the cells it updates, so an interior tile's rings are counted
// geometrically (ring rho of its square) and a step that yields x_k on rho <= lim needs
// x_(k-1) on rho <= lim + 2; a graph tile's rings are BFS rings over the wide stencil graph
// (one ring per step).  The smoothing is the same Chebyshev iteration, of degree D.

// Per-thread view of one ext cell of a Galerkin level.
struct ExtCellW {
	int c;                       // cell index (-1: none)
	int slot;                    // LDS slot
	int ring;
	unsigned nb[GGBK_WNB / 2];   // graph tiles: stencil neighbours' slots, two per word
	unsigned w[GGBK_WNB / 2];    // weights, two half floats per word
	float dinv;
};

__device__ __forceinline__ float half_at(const unsigned *w, int k)
{
	unsigned short h = (unsigned short) (w[k >> 1] >> (16 * (k & 1)));
	return (float) __builtin_bit_cast(_Float16, h);
}

// Sets up ext cell e of a Galerkin level's tile: an interior tile's HS-ring square, or a
// graph tile's list (rings 0..HG); cells with a stencil in the pass (rho <= HS - 2 of a
// square, rings < HG of a graph tile) get their weights (and a graph tile's slots).
template <int HS, int HG>
__device__ __forceinline__ void ext_cell_w(const ggbk_grid &g, const ggbk_wide &wd, const ggbk_tiles &tl, int tile,
					   int kind, int off, int e, int next, const int *ring_end, ExtCellW &x)
{
	x.c = -1;
	if (e >= next)
		return;
	constexpr int W = MGT_T + 2 * HS;
	int c;
	bool st;
	if (kind < 0) {
		int N = g.N, tpe = N / MGT_T;
		int f = tile / (tpe * tpe), r = tile - f * tpe * tpe;
		int J0 = (r / tpe) * MGT_T, I0 = (r % tpe) * MGT_T;
		int ey, ex, ring;
		square_pos<HS>(e, ey, ex, ring);
		c = (f * N + J0 - HS + ey) * N + I0 - HS + ex;
		x.slot = ey * W + ex;
		x.ring = ring;
		st = ring <= HS - 2;
	} else {
		c = ((const int2 *) tl.gcc)[off + e].x;
		x.slot = e;
		int ring = 0;
#pragma unroll
		for (int r = 0; r < HG; r++)
			ring += e >= ring_end[r];
		x.ring = ring;
		st = ring < HG;
		if (st) {
			const uint4 *q = (const uint4 *) (tl.gnbr + GGBK_WNB * (size_t) (off + e));
#pragma unroll
			for (int v = 0; v < 3; v++) {
				uint4 a = q[v];
				x.nb[4 * v + 0] = a.x;
				x.nb[4 * v + 1] = a.y;
				x.nb[4 * v + 2] = a.z;
				x.nb[4 * v + 3] = a.w;
			}
		}
	}
	x.c = c;
	if (st) {
		const uint4 *q = (const uint4 *) (wd.w + GGBK_WNB * (size_t) c);
#pragma unroll
		for (int v = 0; v < 3; v++) {
			uint4 a = q[v];
			x.w[4 * v + 0] = a.x;
			x.w[4 * v + 1] = a.y;
			x.w[4 * v + 2] = a.z;
			x.w[4 * v + 3] = a.w;
		}
	}
	x.dinv = wd.dinv[c];
}

// K x at an ext cell of a Galerkin level from the iterate in LDS, slots in order: an
// interior tile's at the fixed 5x5 offsets (row-major, centre skipped) of its square of
// width W, a graph tile's through its slot list (padded slots: the cell itself, weight 0)
template <int W>
__device__ __forceinline__ float kxw(const ExtCellW &x, const float *xs, bool sq)
{
	float xi = xs[x.slot], s = 0.0f;
	if (sq) {
#pragma unroll
		for (int k = 0; k < GGBK_WNB; k++) {
			int q = k < 12 ? k : k + 1;
			int dy = q / 5 - 2, dx = q % 5 - 2;
			s += half_at(x.w, k) * (xi - xs[x.slot + dy * W + dx]);
		}
	} else {
#pragma unroll
		for (int k = 0; k < GGBK_WNB; k++) {
			int nb = (int) ((x.nb[k >> 1] >> (16 * (k & 1))) & 0xffffu);
			s += half_at(x.w, k) * (xi - xs[nb]);
		}
	}
	return s;
}

// Down pass of a Galerkin level: D Chebyshev steps from 0 on b, the residual on the
// restriction's support and its restriction into the next level (as k_mgt_down).  x_k on
// rho <= 1 + 2 (D + 1 - k) of an interior tile's square (HS = 2 D + 1 rings), on BFS
// rings <= D + 1 - k of a graph tile; the residual on rho <= 1 / ring 0.
template <int D, int PER, int BS = MGT_BS>
__global__ __launch_bounds__(BS) void k_mgw_down(ggbk_grid g, ggbk_wide wd, ggbk_tiles tl, ggbk_xfer xf,
						       const float *b, float *bc, float *xpre, const ggbk_scalars *s)
{
	constexpr int CAP = PER * BS, HS = 2 * D + 1, W = MGT_T + 2 * HS;
	__shared__ float xa[CAP], xb[CAP];
	int act = s->vc;
	if (!act)
		return;
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<HS, D>(g, tl, tile, kind, off, ring_end);
	bool sq = kind < 0;
	int base = sq ? 1 : 0, scale = sq ? 2 : 1;
	ExtCellW xc[PER];
	float bv[PER], dv[PER];
#pragma unroll
	for (int m = 0; m < PER; m++) {
		ext_cell_w<HS, D>(g, wd, tl, tile, kind, off, threadIdx.x + m * BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			bv[m] = b[xc[m].c];
			dv[m] = cheb_cr(0) * xc[m].dinv * bv[m];
			xa[xc[m].slot] = dv[m];
		}
	}
	__syncthreads();
	float *xs = xa, *xn = xb;
#pragma unroll
	for (int k = 1; k < D; k++) {
		int lim = base + scale * (D - k);
#pragma unroll
		for (int m = 0; m < PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= lim) {
				float res = bv[m] - kxw<W>(xc[m], xs, sq);
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
	for (int m = 0; m < PER; m++) {
		if (xc[m].c >= 0 && xc[m].ring <= base) {
			xn[xc[m].slot] = bv[m] - kxw<W>(xc[m], xs, sq);
			bool own = sq ? xc[m].ring == 0 : threadIdx.x + m * BS < MGT_T * MGT_T;
			if (own)
				xpre[xc[m].c] = xs[xc[m].slot];
		}
	}
	__syncthreads();
	restrict_tile<HS>(g, tl, xf, tile, kind, xn, bc, act);
}

// Up pass of a Galerkin level: x = x_pre + P x_coarse, then D Chebyshev steps on b; x_k on
// rho <= 2 (D - k) of an interior tile's square (HS = 2 D rings), on BFS rings <= D - k of
// a graph tile; x_D stored on the tile.
template <int D, int PER, int BS = MGT_BS>
__global__ __launch_bounds__(BS) void k_mgw_up(ggbk_grid g, ggbk_wide wd, ggbk_tiles tl, ggbk_xfer xf,
						     const float *b, const float *xpre, const float *xcoarse, float *xout,
						     const ggbk_scalars *s)
{
	constexpr int CAP = PER * BS, HS = 2 * D, W = MGT_T + 2 * HS;
	__shared__ float xa[CAP], xb[CAP];
	int act = s->vc;
	if (!act)
		return;
	int tile = xblock(), kind, off, ring_end[D + 1];
	int next = tile_setup<HS, D>(g, tl, tile, kind, off, ring_end);
	bool sq = kind < 0;
	int scale = sq ? 2 : 1;
	ExtCellW xc[PER];
	float bv[PER], dv[PER];
#pragma unroll
	for (int m = 0; m < PER; m++) {
		ext_cell_w<HS, D>(g, wd, tl, tile, kind, off, threadIdx.x + m * BS, next, ring_end, xc[m]);
		bv[m] = 0.0f;
		dv[m] = 0.0f;
		if (xc[m].c >= 0) {
			int c = xc[m].c;
			if (xc[m].ring <= scale * (D - 1))
				bv[m] = b[c];
			xa[xc[m].slot] = xpre[c] + prolong_at(xf, xcoarse, c);
		}
	}
	__syncthreads();
	float *xs = xa, *xn = xb;
#pragma unroll
	for (int k = 0; k < D; k++) {
		int lim = scale * (D - 1 - k);
#pragma unroll
		for (int m = 0; m < PER; m++) {
			if (xc[m].c >= 0 && xc[m].ring <= lim) {
				float res = bv[m] - kxw<W>(xc[m], xs, sq);
				dv[m] = cheb_cd(k) * dv[m] + cheb_cr(k) * xc[m].dinv * res;
				float xv = xs[xc[m].slot] + dv[m];
				if (k < D - 1)
					xn[xc[m].slot] = xv;
				else
					xout[xc[m].c] = xv;
			}
		}
		if (k < D - 1) {
			__syncthreads();
			float *t = xs;
			xs = xn;
			xn = t;
		}
	}
}

// Dense bottom (mg_k.hip), gated by s->vc
extern "C" void ggbk_mg_dense_vc_f32(const float *M, int n, const float *b, float *x, const ggbk_scalars *s);

// The Galerkin levels' passes: a workgroup of MGW_BS threads holds a tile's ext cells, one
// per thread (676 in a down pass, 576 in an up pass): measured faster than 256 threads with
// three cells each (whose 110 registers allow 4 waves per SIMD) or 512 with two.
#define MGW_BS 768

// down pass of level L into level C's b: Galerkin levels degree 2, Voronoi levels 3
static void down_level(const ggbk_level &L, const ggbk_level &C, const float *b, ggbk_scalars *s)
{
	if (L.wd.w) {
		GGBK_LAUNCH((k_mgw_down<2, 1, MGW_BS>), L.td.ntiles, MGW_BS, 0, L.g, L.wd, L.td, L.xf, b, (float *) C.b,
			    (float *) L.xpre, s);
		return;
	}
	ggbk_cg_cheb none;
	memset(&none, 0, sizeof(none));
	GGBK_LAUNCH((k_mgt_down<3, float, false, 3>), L.td.ntiles, MGT_BS, 0, L.g, L.td, L.xf, b, (float *) C.b,
		    (float *) L.xpre, none, s);
}

static void up_level(const ggbk_level &L, const float *b, const float *xc, float *x, ggbk_scalars *s)
{
	if (L.wd.w) {
		GGBK_LAUNCH((k_mgw_up<2, 1, MGW_BS>), L.tl.ntiles, MGW_BS, 0, L.g, L.wd, L.tl, L.xf, b,
			    (const float *) L.xpre, xc, x, s);
		return;
	}
	GGBK_LAUNCH((k_mgt_up<3>), L.tl.ntiles, MGT_BS, 0, L.g, L.tl, L.xf, b, (const float *) L.xpre, xc, x, s);
}

// level 0's passes for its smoothing degree D (2 ext cells per thread for D = 2, else 3;
// 384 or 512 threads with two cells each measured slower)
template <int D, int PER, int BS = MGT_BS>
static void level0_down(const ggbk_level *lv, const ggbk_cg_cheb &cg, ggbk_scalars *s)
{
	const ggbk_level &L0 = lv[0];
	if (cg.update)
		GGBK_LAUNCH((k_mgt_down<D, double, true, PER, BS>), L0.td.ntiles, BS, 0, L0.g, L0.td, L0.xf,
			    (const double *) nullptr, (float *) lv[1].b, (float *) L0.xpre, cg, s);
	else
		GGBK_LAUNCH((k_mgt_down<D, double, false, PER, BS>), L0.td.ntiles, BS, 0, L0.g, L0.td, L0.xf, cg.r,
			    (float *) lv[1].b, (float *) L0.xpre, cg, s);
}

template <int D, int PER, int BS = MGT_BS>
static void level0_up(const ggbk_level *lv, const double *b0, void *zout, double *Az, const double *ap,
		      ggbk_scalars *s, ggbk_red red)
{
	const ggbk_level &L0 = lv[0];
	GGBK_LAUNCH((k_mgt_up_cg<D, PER, BS>), L0.tl.ntiles, BS, 0, L0.g, L0.tl, L0.xf, b0, (const float *) L0.xpre,
		    (const float *) lv[1].x, (float *) zout, Az, ap, s, red);
}

extern "C" void ggbk_cg_update_cheb(ggbk_grid g, const ggbk_cg_cheb *c, const ggbk_scalars *s)
{
	ggbk_probe_begin(GGBK_PROBE_CG_UPDATE);
	GGBK_LAUNCH(k_cg_update_cheb, (unsigned) ((g.n + MGT_BS - 1) / MGT_BS), MGT_BS, 0, g.n, *c, s);
	ggbk_probe_end(GGBK_PROBE_CG_UPDATE);
}

extern "C" void ggbk_vcycle_cheb(const ggbk_level *lv, int nlev, const ggbk_cg_cheb *c, const void *r, void *zout,
				 ggbk_scalars *s, ggbk_red red)
{
	(void) nlev;
	const ggbk_level &L0 = lv[0];
	ggbk_cg_cheb cg;
	memset(&cg, 0, sizeof(cg));
	if (c)
		cg = *c;
	else
		cg.r = (const double *) r;
	// level 0 down into level 1, fused with the PCG update when asked
	int D0 = L0.td.deg;
	ggbk_probe_begin(cg.update ? GGBK_PROBE_CG_UPDATE : GGBK_PROBE_MG_DOWN0);
	if (D0 == 2)
		level0_down<2, 2>(lv, cg, s);
	else if (D0 == 3)
		level0_down<3, 3>(lv, cg, s);
	else
		level0_down<4, 3>(lv, cg, s);
	ggbk_probe_end(cg.update ? GGBK_PROBE_CG_UPDATE : GGBK_PROBE_MG_DOWN0);
	const double *b0 = cg.update ? cg.r_out : cg.r;
	// smooth and restrict down to the dense level
	int l = 1;
	while (!lv[l].M) {
		down_level(lv[l], lv[l + 1], (const float *) lv[l].b, s);
		l++;
	}
	ggbk_mg_dense_vc_f32((const float *) lv[l].M, lv[l].g.n, (const float *) lv[l].b, (float *) lv[l].x, s);
	// back up
	for (l = l - 1; l >= 1; l--) {
		const ggbk_level &L = lv[l];
		up_level(L, (const float *) L.b, (const float *) lv[l + 1].x, (float *) L.x, s);
	}
	// the current direction's image: just formed by the update, else the caller's
	const double *ap = cg.update ? cg.Ap_out : cg.Ap_in;
	double *Az = c ? cg.Az : nullptr;
	ggbk_red rd = c ? red : ggbk_red{ nullptr, 0, nullptr };
	ggbk_probe_begin(GGBK_PROBE_MG_UP0);
	if (D0 == 2)
		level0_up<2, 2>(lv, b0, zout, Az, ap, s, rd);
	else if (D0 == 3)
		level0_up<3, 3>(lv, b0, zout, Az, ap, s, rd);
	else
		level0_up<4, 3>(lv, b0, zout, Az, ap, s, rd);
	ggbk_probe_end(GGBK_PROBE_MG_UP0);
	// the fold (after the first V-cycle of a solve there is no direction yet)
	if (c)
		ggbk_fold_cheb(s, red, L0.tl.ntiles, !cg.update && !cg.Ap_in, cg.start ? cg.nb0 : 0, cg.tol, cg.maxit);
}
