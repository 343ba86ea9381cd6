// dev.h -- device-side math shared by the gfx950 kernels: double 3-vectors, the
// great-circle step, cubed-sphere location and bilinear taps, and deterministic
// block reductions (wave64 shuffles, then LDS, fixed order).
//
// The vector helpers keep the reference's operation order (src/vec3.h:9-42; e.g.
// normalise = multiply by the reciprocal length) and every kernel is compiled with
// -ffp-contract=off, so the only source of difference from the CPU path is the device
// libm (atan, atan2, sin, cos), which is not bit-identical to glibc.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ggbk.h"

#define GGB_PI 3.14159265358979323846

struct dv3 {
	double x, y, z;
};

__device__ __forceinline__ dv3 d3(double x, double y, double z)
{
	dv3 r;
	r.x = x;
	r.y = y;
	r.z = z;
	return r;
}
__device__ __forceinline__ dv3 vadd(dv3 a, dv3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ dv3 vsub(dv3 a, dv3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ dv3 vscl(dv3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double vdot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ dv3 vcross(dv3 a, dv3 b)
{
	return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double vlen(dv3 a) { return sqrt(vdot(a, a)); }
__device__ __forceinline__ dv3 vnorm(dv3 a)
{
	double l = vlen(a);
	return l > 0.0 ? vscl(a, 1.0 / l) : a;
}
__device__ __forceinline__ dv3 vtan(dv3 v, dv3 n) { return vsub(v, vscl(n, vdot(v, n))); }
__device__ __forceinline__ double vangle(dv3 a, dv3 b) { return atan2(vlen(vcross(a, b)), vdot(a, b)); }

template <typename T>
__device__ __forceinline__ dv3 ld3(const T *p)
{
	return d3((double) p[0], (double) p[1], (double) p[2]);
}
template <typename T>
__device__ __forceinline__ void st3(T *p, dv3 a)
{
	p[0] = (T) a.x;
	p[1] = (T) a.y;
	p[2] = (T) a.z;
}

// Great-circle move (reference vec3.h:33-42).
__device__ __forceinline__ dv3 gc_move(dv3 p, dv3 v, double s)
{
	dv3 t = vtan(v, p);
	double sp = vlen(t);
	if (sp < 1e-300)
		return p;
	double th = sp * s;
	double sn, cs;
	sincos(th, &sn, &cs);
	return vnorm(vadd(vscl(p, cs), vscl(t, sn / sp)));
}

// Face of direction d: largest |component|, ties to x then y (reference cubesphere.c:12-20).
__device__ __forceinline__ int face_of(dv3 d)
{
	double ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
	if (ax >= ay && ax >= az)
		return d.x > 0 ? 1 : 3;
	if (ay >= az)
		return d.y > 0 ? 4 : 5;
	return d.z > 0 ? 0 : 2;
}

// Gnomonic coordinates on face f: dc = d.centre, u = d.right / dc, v = d.up / dc.  The
// face basis vectors have one +-1 entry, so the dot products are single components
// (bitwise equal to the full dot products for dc != 0).
__device__ __forceinline__ void gnomonic(int f, dv3 d, double &dc, double &u, double &v)
{
	double cr, cu;
	switch (f) {
	case 0: dc = d.z; cr = d.x; cu = d.y; break;
	case 1: dc = d.x; cr = -d.z; cu = d.y; break;
	case 2: dc = -d.z; cr = -d.x; cu = d.y; break;
	case 3: dc = -d.x; cr = d.z; cu = d.y; break;
	case 4: dc = d.y; cr = d.x; cu = -d.z; break;
	default: dc = -d.y; cr = d.x; cu = d.z; break;
	}
	u = cr / dc;
	v = cu / dc;
}

// Dot product of d with the centre of face f (any f; used by the splat face test).
__device__ __forceinline__ double face_dot(int f, dv3 d)
{
	switch (f) {
	case 0: return d.z;
	case 1: return d.x;
	case 2: return -d.z;
	case 3: return -d.x;
	case 4: return d.y;
	default: return -d.y;
	}
}

// Continuous cell coordinates (reference cubesphere.c:29-38).
__device__ __forceinline__ void locate(double h, dv3 d, int &f, double &x, double &y)
{
	f = face_of(d);
	double dc, u, v;
	gnomonic(f, d, dc, u, v);
	x = (atan(u) + GGB_PI / 4) / h - 0.5;
	y = (atan(v) + GGB_PI / 4) / h - 0.5;
}

// Four bilinear taps at p (reference cubesphere.c:68-92).  Cells off the face resolve
// through the ghost table built on the host with the reference's rule.
struct Taps {
	int idx[4];
	double fx, fy;
};

__device__ __forceinline__ Taps taps_at(const ggbk_grid &g, dv3 p)
{
	Taps t;
	int f;
	double x, y;
	locate(g.h, p, f, x, y);
	int N = g.N;
	// x, y lie in [-0.5, N - 0.5] for any unit vector; the clamp only keeps a
	// non-finite position from indexing outside the ghost table
	double xf = fmin(fmax(floor(x), -1.0), (double) (N - 1));
	double yf = fmin(fmax(floor(y), -1.0), (double) (N - 1));
	int i0 = (int) xf, j0 = (int) yf;
	t.fx = x - i0;
	t.fy = y - j0;
	if (i0 >= 0 && j0 >= 0 && i0 + 1 < N && j0 + 1 < N) {
		int b = (f * N + j0) * N + i0;
		t.idx[0] = b;
		t.idx[1] = b + 1;
		t.idx[2] = b + N;
		t.idx[3] = b + N + 1;
	} else {
		int P = N + 2;
		const int *gt = g.ghost + (size_t) f * P * P;
		int r0 = (j0 + 1) * P + (i0 + 1);
		t.idx[0] = gt[r0];
		t.idx[1] = gt[r0 + 1];
		t.idx[2] = gt[r0 + P];
		t.idx[3] = gt[r0 + P + 1];
	}
	return t;
}

__device__ __forceinline__ void tap_weights(const Taps &t, double wt[4])
{
	wt[0] = (1 - t.fx) * (1 - t.fy);
	wt[1] = t.fx * (1 - t.fy);
	wt[2] = (1 - t.fx) * t.fy;
	wt[3] = t.fx * t.fy;
}

template <typename T>
__device__ __forceinline__ double sample_scalar(const ggbk_grid &g, const T *fld, dv3 p)
{
	Taps t = taps_at(g, p);
	double wt[4];
	tap_weights(t, wt);
	return wt[0] * (double) fld[t.idx[0]] + wt[1] * (double) fld[t.idx[1]] +
	       wt[2] * (double) fld[t.idx[2]] + wt[3] * (double) fld[t.idx[3]];
}

template <typename T>
__device__ __forceinline__ dv3 sample_vec(const ggbk_grid &g, const T *fld3, dv3 p)
{
	Taps t = taps_at(g, p);
	double wt[4];
	tap_weights(t, wt);
	dv3 r = d3(0, 0, 0);
	for (int k = 0; k < 4; k++)
		r = vadd(r, vscl(ld3(fld3 + 3 * (size_t) t.idx[k]), wt[k]));
	return vtan(r, p);
}

__device__ __forceinline__ dv3 cell_pos(const ggbk_grid &g, int i) { return ld3(g.pos + 3 * (size_t) i); }

// Visits the Laplacian stencil of cell i in the reference's slot order: fn(j, w) for each
// neighbour j with weight w.  Slots 0-5 are always visited (cells with fewer neighbours
// are padded with (i, 0), which add exact zeros after the real terms), slots 6-7 only
// where the cell has them.  The cell's record (8 bytes) gives its weight class, whose
// row of 8 weights is shared by the ~47 cells related by the cube's symmetries and is
// read from cache, and its neighbours as 4-bit offset codes (or, for cells on face seams,
// the index of an explicit row).
template <typename Fn>
__device__ __forceinline__ void for_stencil(const ggbk_grid &g, int i, Fn fn)
{
	int2 rc = ((const int2 *) g.srec)[i];
	const float4 *wp = (const float4 *) (g.scw + 8 * (size_t) (rc.x & 0x7fffffff));
	float4 wa = wp[0], wb = wp[1];
	int nb[8];
	if (rc.x < 0) {
		const int4 *xp = (const int4 *) (g.sxn + 8 * (size_t) rc.y);
		int4 a = xp[0], b = xp[1];
		nb[0] = a.x; nb[1] = a.y; nb[2] = a.z; nb[3] = a.w;
		nb[4] = b.x; nb[5] = b.y; nb[6] = b.z; nb[7] = b.w;
	} else {
		// slot k: code (dy + 1) << 2 | (dx + 1), neighbour i + dy N + dx
		unsigned code = (unsigned) rc.y;
		int base = i - g.N - 1;
#pragma unroll
		for (int k = 0; k < 8; k++)
			nb[k] = base + (int) ((code >> (4 * k + 2)) & 3u) * g.N + (int) ((code >> (4 * k)) & 3u);
	}
	fn(nb[0], wa.x);
	fn(nb[1], wa.y);
	fn(nb[2], wa.z);
	fn(nb[3], wa.w);
	fn(nb[4], wb.x);
	fn(nb[5], wb.y);
	if (g.maxnb > 6) {
		if (nb[6] != i)
			fn(nb[6], wb.z);
		if (nb[7] != i)
			fn(nb[7], wb.w);
	}
}

// ---------------- deterministic reductions ----------------

// Wave64 tree sum via DPP-free shuffles; every lane returns the wave total.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v = fmax(v, __shfl_xor(v, off, 64));
	return v;
}

// Block sum in a fixed order: wave totals, then wave 0 folds them.  Result valid in
// thread 0.  BS = block size (multiple of 64, <= 1024).
template <int BS>
__device__ __forceinline__ double block_sum(double v)
{
	__shared__ double wsum[BS / 64];
	v = wave_sum(v);
	int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (lane == 0)
		wsum[wid] = v;
	__syncthreads();
	double r = 0;
	if (threadIdx.x == 0) {
		for (int k = 0; k < BS / 64; k++)
			r += wsum[k];
	}
	__syncthreads();
	return r;
}

template <int BS>
__device__ __forceinline__ double block_max(double v)
{
	__shared__ double wmax[BS / 64];
	v = wave_max(v);
	int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (lane == 0)
		wmax[wid] = v;
	__syncthreads();
	double r = 0;
	if (threadIdx.x == 0) {
		for (int k = 0; k < BS / 64; k++)
			r = fmax(r, wmax[k]);
	}
	__syncthreads();
	return r;
}

// The mixed-precision PCG's scalar update after a V-cycle, from its nine sums v = r.z,
// z.Az, z.Ap, r.Az, Az.Az, Az.Ap, r.Ap, Ap.Ap, r.r (Ap the current direction's image): rz,
// beta = rz / rz_old, p.Ap = z.Az + 2 beta z.Ap + beta^2 p.Ap_old (A symmetric), alpha, and the
// residual after the coming update, r.r - 2 alpha r.Ap' + alpha^2 Ap'.Ap' with
// Ap' = Az + beta Ap; the coming update is the last (stop) when that meets the tolerance
// or the iteration cap.  first: no previous direction (p = z, Ap = Az).  r.r is the sum
// over r itself: the prediction loses about two digits to cancellation, and chaining
// predictions would compound that each iteration.
__device__ __forceinline__ void ggbk_cheb_fold(ggbk_scalars *s, const double *v, int first)
{
	if (!s->active)
		return;
	if (s->stop) { // the last update has been applied
		s->active = 0;
		s->vc = 0;
		return;
	}
	double rz = v[0], zAz = v[1], zAp = v[2], rAz = v[3], AzAz = v[4], AzAp = v[5], rAp0 = v[6], ApAp0 = v[7];
	double rr = v[8], beta, pAp, rAp, ApAp;
	if (first) {
		beta = 0.0;
		pAp = zAz;
		rAp = rAz;
		ApAp = AzAz;
	} else {
		beta = rz / s->rz;
		pAp = zAz + 2.0 * beta * zAp + beta * beta * s->pAp;
		rAp = rAz + beta * rAp0;
		ApAp = AzAz + 2.0 * beta * AzAp + beta * beta * ApAp0;
	}
	s->rz = rz;
	s->beta = beta;
	s->pAp = pAp;
	if (!(pAp > 0)) {
		s->status = 1;
		s->active = 0;
		s->vc = 0;
		return;
	}
	double alpha = rz / pAp;
	s->alpha = alpha;
	double rrn = rr - 2.0 * alpha * rAp + alpha * alpha * ApAp;
	rrn = rrn > 0.0 ? rrn : 0.0;
	s->rr = rrn;
	s->it = s->it + 1;
	if (rrn <= s->tol2) {
		s->stop = 1;
		s->vc = 0;
		s->status = 0;
	} else if (s->it >= s->maxit) {
		s->stop = 1;
		s->vc = 0;
		s->status = 3;
	}
}

// Logical block index with each XCD's share of the grid contiguous.  The dispatcher
// deals workgroups to the 8 XCDs round-robin (workgroup k runs on XCD k mod 8) and each
// XCD has its own L2, so with the plain index the cells of rows i +- N, which a stencil
// gathers, are fetched by three L2s.  Here XCD x runs logical blocks
// [x q + min(x, r), ...) with q = nb / 8, r = nb % 8: a bijection onto [0, nb).
#define GGBK_NXCD 8
__device__ __forceinline__ int xblock(void)
{
	int k = blockIdx.x, nb = gridDim.x;
	int x = k % GGBK_NXCD, idx = k / GGBK_NXCD;
	int q = nb / GGBK_NXCD, r = nb % GGBK_NXCD;
	return x * q + min(x, r) + idx;
}

// ---------------- face tiles staged in LDS ----------------
// A workgroup of BS = 256 threads takes a 16x16 tile of one face (cell (f, J, I) with J, I
// in the tile) and stages a per-cell value of the tile and its one-cell ring (18x18, the
// ring only where it lies on the face) in LDS.  A cell whose stencil record holds offset
// codes (all neighbours on its own face, offsets dy N + dx) then reads its neighbours'
// values from LDS at (ty + dy, tx + dx); a cell on a face seam (explicit neighbour row)
// computes them from global memory.  Needs N divisible by 16 and workgroups of 256 threads.
#define TT 16
#define TH (TT + 2)

struct FaceTile {
	int f, J0, I0, ty, tx, c;
};

__device__ __forceinline__ FaceTile face_tile(int N)
{
	FaceTile t;
	int tpf = (N / TT) * (N / TT);
	int b = xblock();
	t.f = b / tpf;
	int r = b - t.f * tpf;
	t.J0 = (r / (N / TT)) * TT;
	t.I0 = (r % (N / TT)) * TT;
	t.ty = threadIdx.x / TT;
	t.tx = threadIdx.x % TT;
	t.c = (t.f * N + t.J0 + t.ty) * N + t.I0 + t.tx;
	return t;
}

// Calls val(cell) for the tile's ring cells on the face and stores the results at their
// LDS places; interior values are stored by the caller at (ty + 1, tx + 1).
template <typename V, typename Val>
__device__ __forceinline__ void stage_ring(const FaceTile &t, int N, V *lds, Val val)
{
	int h = threadIdx.x;
	if (h < 4 * TH - 4) {
		int hy, hx;
		if (h < TH) {
			hy = -1;
			hx = h - 1;
		} else if (h < 2 * TH) {
			hy = TT;
			hx = h - TH - 1;
		} else if (h < 2 * TH + TT) {
			hy = h - 2 * TH;
			hx = -1;
		} else {
			hy = h - 2 * TH - TT;
			hx = TT;
		}
		int J = t.J0 + hy, I = t.I0 + hx;
		if (J >= 0 && J < N && I >= 0 && I < N)
			lds[(hy + 1) * TH + hx + 1] = val((t.f * N + J) * N + I);
	}
}

// sum_k w_k (xi - v_k) over cell t.c's stencil in slot order, the neighbour values v from
// the staged tile (cells with offset codes) or from gval(j) (seam cells); the same terms,
// in the same order, as for_stencil with v = gval(j).  Call after the LDS is staged.
template <typename V, typename GVal>
__device__ __forceinline__ double tile_stencil_sum(const ggbk_grid &g, const FaceTile &t, const V *lds, double xi,
						   GVal gval)
{
	double sum = 0;
	int2 rc = ((const int2 *) g.srec)[t.c];
	if (rc.x >= 0) {
		const float4 *wp = (const float4 *) (g.scw + 8 * (size_t) rc.x);
		float4 wa = wp[0], wb = wp[1];
		float w[8] = { wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w };
		unsigned code = (unsigned) rc.y;
		int l0 = (t.ty + 1) * TH + t.tx + 1;
#pragma unroll
		for (int k = 0; k < 8; k++) {
			int dy = (int) ((code >> (4 * k + 2)) & 3u) - 1, dx = (int) ((code >> (4 * k)) & 3u) - 1;
			if (k < 6 || (dy | dx) != 0)
				sum += (double) w[k] * (xi - (double) lds[l0 + dy * TH + dx]);
		}
	} else {
		for_stencil(g, t.c, [&](int j, float w) { sum += (double) w * (xi - (double) gval(j)); });
	}
	return sum;
}

// Kernel launch error check helper for host launchers.
#define GGBK_LAUNCH(kern, grid, block, shm, ...)                                           \
	hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, ggbk_stream(), __VA_ARGS__)

hipStream_t ggbk_stream(void);

static inline unsigned int ggbk_blocks(long long n, int bs)
{
	long long b = (n + bs - 1) / bs;
	return (unsigned int) (b < 1 ? 1 : b);
}
