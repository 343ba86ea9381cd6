// render_k.hip -- deterministic trail splatting (reference src/render.c:87-218).
//
// The reference emits one record per (sample, face, 32-row strip), sorts them
// bin-major / thread-minor and blends each bin in order, so every texel sees its
// splats in (particle index, sample) order.  Here the unit of work is a face TILE of
// TILE x TILE texels held in LDS:
//   1. count  -- each particle walks its arc samples and adds itself once to every
//                tile its footprints' bounding box (per face) overlaps;
//   2. scan   -- exclusive scan of the tile counts (one workgroup);
//   3. fill   -- the same walk writes the particle index into each tile's list;
//   4. draw   -- one workgroup per tile: sort the list (bitonic, LDS), load the tile,
//                then blend particles in ascending index.  Particles whose footprint
//                boxes overlap are serialised by an LDS ownership pass (atomicMin of
//                the window rank per texel): a particle blends only when it is the
//                lowest unprocessed claimant of every texel in its box, so each texel
//                sees the reference order.  Each particle blends its samples in k order.
// The texel arithmetic is the reference's (render.c:113-154) in float with no FMA
// contraction, so texels match the CPU path bit for bit for identical positions.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "dev.h"
#include "ggbk.h"

#define PBS 256          // particle kernels
#define DBS 256          // draw workgroup
#define TILE_MAX 32
#define SORT_CAP 2048    // LDS sort capacity per tile
#define SCAN_BS 1024

struct SplatGeo {
	int R;
	float rad;
	double spacing;
};

__device__ __forceinline__ int arc_samples(dv3 a, dv3 b, double spacing)
{
	double ang = vangle(a, b);
	int ns = (int) ceil(ang / spacing);
	if (ns < 1)
		ns = 1;
	if (ns > 4096)
		ns = 4096;
	return ns;
}

__device__ __forceinline__ dv3 arc_point(dv3 a, dv3 b, int k, int ns)
{
	double tt = (double) k / ns;
	return vnorm(vadd(vscl(a, 1 - tt), vscl(b, tt)));
}

// Pixel position of direction d on face f (render.c:91-98); false when the face is
// not within 60 degrees (dc < 0.5).
__device__ __forceinline__ bool face_pixel(int f, dv3 d, int R, float &px, float &py)
{
	double dc, u, v;
	gnomonic(f, d, dc, u, v);
	if (!(face_dot(f, d) >= 0.5))
		return false;
	px = (float) ((u + 1.0) * 0.5 * R - 0.5);
	py = (float) ((1.0 - v) * 0.5 * R - 0.5);
	return true;
}

// Footprint box clipped to the face; false if empty (render.c:99-109).
__device__ __forceinline__ bool footprint(float px, float py, float rad, int R, int &xlo, int &xhi, int &ylo, int &yhi)
{
	xlo = (int) ceilf(px - rad);
	xhi = (int) floorf(px + rad);
	ylo = (int) ceilf(py - rad);
	yhi = (int) floorf(py + rad);
	if (xlo < 0) xlo = 0;
	if (xhi > R - 1) xhi = R - 1;
	if (ylo < 0) ylo = 0;
	if (yhi > R - 1) yhi = R - 1;
	return xlo <= xhi && ylo <= yhi;
}

// Per-face union of the footprint boxes of a particle's samples.
struct FaceBoxes {
	int xlo[6], xhi[6], ylo[6], yhi[6];
};

__device__ __forceinline__ void arc_boxes(dv3 a, dv3 b, const SplatGeo &sg, FaceBoxes &fb)
{
	for (int f = 0; f < 6; f++) {
		fb.xlo[f] = INT_MAX;
		fb.ylo[f] = INT_MAX;
		fb.xhi[f] = INT_MIN;
		fb.yhi[f] = INT_MIN;
	}
	int ns = arc_samples(a, b, sg.spacing);
	for (int k = 1; k <= ns; k++) {
		dv3 d = arc_point(a, b, k, ns);
		for (int f = 0; f < 6; f++) {
			if (!(face_dot(f, d) >= 0.5))
				continue;
			float px, py;
			face_pixel(f, d, sg.R, px, py);
			int xlo, xhi, ylo, yhi;
			if (!footprint(px, py, sg.rad, sg.R, xlo, xhi, ylo, yhi))
				continue;
			fb.xlo[f] = min(fb.xlo[f], xlo);
			fb.xhi[f] = max(fb.xhi[f], xhi);
			fb.ylo[f] = min(fb.ylo[f], ylo);
			fb.yhi[f] = max(fb.yhi[f], yhi);
		}
	}
}

__global__ __launch_bounds__(PBS) void k_render_count(int np, const double *prev, const double *pos, SplatGeo sg,
						      int tile, int ntx, int *counts)
{
	int i = blockIdx.x * PBS + threadIdx.x;
	if (i >= np)
		return;
	FaceBoxes fb;
	arc_boxes(ld3(prev + 3 * (size_t) i), ld3(pos + 3 * (size_t) i), sg, fb);
	for (int f = 0; f < 6; f++) {
		if (fb.xlo[f] > fb.xhi[f])
			continue;
		for (int ty = fb.ylo[f] / tile; ty <= fb.yhi[f] / tile; ty++)
			for (int tx = fb.xlo[f] / tile; tx <= fb.xhi[f] / tile; tx++)
				atomicAdd(&counts[(f * ntx + ty) * ntx + tx], 1);
	}
}

__global__ __launch_bounds__(PBS) void k_render_fill(int np, const double *prev, const double *pos, SplatGeo sg,
						     int tile, int ntx, const int *starts, int *cursor, int *list,
						     long long cap)
{
	int i = blockIdx.x * PBS + threadIdx.x;
	if (i >= np)
		return;
	FaceBoxes fb;
	arc_boxes(ld3(prev + 3 * (size_t) i), ld3(pos + 3 * (size_t) i), sg, fb);
	for (int f = 0; f < 6; f++) {
		if (fb.xlo[f] > fb.xhi[f])
			continue;
		for (int ty = fb.ylo[f] / tile; ty <= fb.yhi[f] / tile; ty++)
			for (int tx = fb.xlo[f] / tile; tx <= fb.xhi[f] / tile; tx++) {
				int t = (f * ntx + ty) * ntx + tx;
				long long slot = (long long) starts[t] + atomicAdd(&cursor[t], 1);
				if (slot < cap)
					list[slot] = i;
			}
	}
}

// Exclusive scan of tile counts (one workgroup).  stats[0] = total entries,
// stats[1] = longest tile list.
__global__ __launch_bounds__(SCAN_BS) void k_render_scan(int ntiles, const int *counts, int *starts, int *cursor,
							 long long *stats)
{
	__shared__ long long part[SCAN_BS];
	__shared__ int pmax[SCAN_BS];
	int per = (ntiles + SCAN_BS - 1) / SCAN_BS;
	int lo = threadIdx.x * per, hi = min(lo + per, ntiles);
	long long s = 0;
	int m = 0;
	for (int t = lo; t < hi; t++) {
		s += counts[t];
		m = max(m, counts[t]);
	}
	part[threadIdx.x] = s;
	pmax[threadIdx.x] = m;
	__syncthreads();
	// Hillis-Steele inclusive scan of the per-thread sums
	for (int off = 1; off < SCAN_BS; off <<= 1) {
		long long v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
		int mv = threadIdx.x >= off ? pmax[threadIdx.x - off] : 0;
		__syncthreads();
		part[threadIdx.x] += v;
		pmax[threadIdx.x] = max(pmax[threadIdx.x], mv);
		__syncthreads();
	}
	long long acc = threadIdx.x ? part[threadIdx.x - 1] : 0;
	for (int t = lo; t < hi; t++) {
		starts[t] = (int) acc;
		acc += counts[t];
		cursor[t] = 0;
	}
	if (threadIdx.x == SCAN_BS - 1) {
		starts[ntiles] = (int) part[SCAN_BS - 1];
		stats[0] = part[SCAN_BS - 1];
		stats[1] = pmax[SCAN_BS - 1];
	}
}

__device__ __forceinline__ void bitonic_sort(int *a, int P)
{
	for (int k = 2; k <= P; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < P; i += DBS) {
				int ixj = i ^ j;
				if (ixj > i) {
					int x = a[i], y = a[ixj];
					bool up = (i & k) == 0;
					if ((x > y) == up) {
						a[i] = y;
						a[ixj] = x;
					}
				}
			}
			__syncthreads();
		}
}

struct TileRect {
	int f, x0, y0, w, h;
};

// Blend all samples of one particle that fall on this tile (render.c:113-154 applied
// to the tile's texels).
__device__ void blend_particle(float4 *tex, const TileRect &tr, dv3 a, dv3 b, const float *colp, const SplatGeo &sg,
			       float alpha, float invg)
{
	float rad = sg.rad, invrad = 1.0f / rad;
	float c0 = colp[0], c1 = colp[1], c2 = colp[2], c3 = colp[3];
	float w0 = alpha * c3;
	int ns = arc_samples(a, b, sg.spacing);
	for (int k = 1; k <= ns; k++) {
		dv3 d = arc_point(a, b, k, ns);
		float px, py;
		if (!face_pixel(tr.f, d, sg.R, px, py))
			continue;
		int xlo, xhi, ylo, yhi;
		if (!footprint(px, py, rad, sg.R, xlo, xhi, ylo, yhi))
			continue;
		xlo = max(xlo, tr.x0);
		xhi = min(xhi, tr.x0 + tr.w - 1);
		ylo = max(ylo, tr.y0);
		yhi = min(yhi, tr.y0 + tr.h - 1);
		for (int y = ylo; y <= yhi; y++) {
			float wy = 1.0f - fabsf((float) y - py) * invrad;
			if (wy <= 0)
				continue;
			for (int x = xlo; x <= xhi; x++) {
				float wx = 1.0f - fabsf((float) x - px) * invrad;
				if (wx <= 0)
					continue;
				float w = w0 * wx * wy;
				float keep = 1.0f - w, add = w * invg;
				float4 &dd = tex[(y - tr.y0) * TILE_MAX + (x - tr.x0)];
				dd.x = dd.x * keep + add * c0;
				dd.y = dd.y * keep + add * c1;
				dd.z = dd.z * keep + add * c2;
				dd.w = dd.w * keep + add;
			}
		}
	}
}

// Box of the particle's footprints on this tile (false if none).
__device__ bool tile_box(const TileRect &tr, dv3 a, dv3 b, const SplatGeo &sg, int &bx0, int &bx1, int &by0, int &by1)
{
	bx0 = INT_MAX;
	by0 = INT_MAX;
	bx1 = INT_MIN;
	by1 = INT_MIN;
	int ns = arc_samples(a, b, sg.spacing);
	for (int k = 1; k <= ns; k++) {
		dv3 d = arc_point(a, b, k, ns);
		float px, py;
		if (!face_pixel(tr.f, d, sg.R, px, py))
			continue;
		int xlo, xhi, ylo, yhi;
		if (!footprint(px, py, sg.rad, sg.R, xlo, xhi, ylo, yhi))
			continue;
		xlo = max(xlo, tr.x0);
		xhi = min(xhi, tr.x0 + tr.w - 1);
		ylo = max(ylo, tr.y0);
		yhi = min(yhi, tr.y0 + tr.h - 1);
		if (xlo > xhi || ylo > yhi)
			continue;
		bx0 = min(bx0, xlo);
		bx1 = max(bx1, xhi);
		by0 = min(by0, ylo);
		by1 = max(by1, yhi);
	}
	return bx0 <= bx1;
}

__global__ __launch_bounds__(DBS) void k_render_draw(ggbk_render r, const double *prev, const double *pos,
						     const float *col, SplatGeo sg, float invg, const long long *stats)
{
	__shared__ float4 tex[TILE_MAX * TILE_MAX];
	__shared__ int owner[TILE_MAX * TILE_MAX];
	__shared__ int sbuf[SORT_CAP];
	if (stats[0] > r.cap)
		return; // list overflowed: the host re-runs with a larger list
	int t = blockIdx.x;
	int L = r.tile_start[t + 1] - r.tile_start[t];
	if (L == 0)
		return;
	const int T = r.tile;
	TileRect tr;
	tr.f = t / (r.ntx * r.ntx);
	int rem = t % (r.ntx * r.ntx);
	tr.y0 = (rem / r.ntx) * T;
	tr.x0 = (rem % r.ntx) * T;
	tr.w = min(T, r.R - tr.x0);
	tr.h = min(T, r.R - tr.y0);
	const int *lst = r.list + r.tile_start[t];

	// sorted particle list: in LDS when it fits, else in this tile's scratch region
	int P = 1;
	while (P < L)
		P <<= 1;
	int *sorted;
	if (P <= SORT_CAP) {
		sorted = sbuf;
	} else {
		sorted = r.sort_scratch + 2 * (size_t) r.tile_start[t];
	}
	for (int i = threadIdx.x; i < P; i += DBS)
		sorted[i] = i < L ? lst[i] : INT_MAX;
	__syncthreads();
	bitonic_sort(sorted, P);

	// load the tile
	float4 *gbuf = (float4 *) r.buf;
	for (int i = threadIdx.x; i < TILE_MAX * TILE_MAX; i += DBS) {
		int ly = i / TILE_MAX, lx = i % TILE_MAX;
		owner[i] = INT_MAX;
		if (ly < tr.h && lx < tr.w)
			tex[i] = gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx];
	}
	__syncthreads();

	for (int base = 0; base < L; base += DBS) {
		int slot = base + (int) threadIdx.x;
		int pid = slot < L ? sorted[slot] : -1;
		dv3 a = d3(0, 0, 0), b = d3(0, 0, 0);
		int bx0 = 0, bx1 = -1, by0 = 0, by1 = -1;
		bool pending = false;
		if (pid >= 0) {
			a = ld3(prev + 3 * (size_t) pid);
			b = ld3(pos + 3 * (size_t) pid);
			pending = tile_box(tr, a, b, sg, bx0, bx1, by0, by1);
		}
		int rank = threadIdx.x;
		while (__syncthreads_or(pending)) {
			bool claimed = pending;
			if (claimed)
				for (int y = by0; y <= by1; y++)
					for (int x = bx0; x <= bx1; x++)
						atomicMin(&owner[(y - tr.y0) * TILE_MAX + (x - tr.x0)], rank);
			__syncthreads();
			bool go = claimed;
			if (claimed)
				for (int y = by0; y <= by1 && go; y++)
					for (int x = bx0; x <= bx1 && go; x++)
						go = owner[(y - tr.y0) * TILE_MAX + (x - tr.x0)] == rank;
			__syncthreads();
			if (go) {
				blend_particle(tex, tr, a, b, col + 4 * (size_t) pid, sg, r.alpha, invg);
				pending = false;
			}
			__syncthreads();
			if (claimed)
				for (int y = by0; y <= by1; y++)
					for (int x = bx0; x <= bx1; x++)
						owner[(y - tr.y0) * TILE_MAX + (x - tr.x0)] = INT_MAX;
			__syncthreads();
		}
	}

	for (int i = threadIdx.x; i < TILE_MAX * TILE_MAX; i += DBS) {
		int ly = i / TILE_MAX, lx = i % TILE_MAX;
		if (ly < tr.h && lx < tr.w)
			gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx] = tex[i];
	}
}

__global__ void k_render_scale(float *buf, size_t count, float g)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t) gridDim.x * blockDim.x;
	for (; i < count; i += stride)
		buf[i] *= g;
}

extern "C" void ggbk_render_scale(float *buf, size_t count, float g)
{
	long long blocks = (long long) ((count + 255) / 256);
	if (blocks > 65536)
		blocks = 65536;
	GGBK_LAUNCH(k_render_scale, (unsigned) blocks, 256, 0, buf, count, g);
}

// Host side of one deposit.  Returns the number of (particle, tile) entries, or -1.
// The only synchronisation is the read-back of the entry total after the scan
// (the list is grown when the total exceeds its capacity).
extern "C" long long ggbk_render_deposit(ggbk_render *r, int np, const double *prev, const double *pos,
					 const float *col, float invg)
{
	if (np <= 0)
		return 0;
	if (r->tile > TILE_MAX || r->tile < 1)
		return -1;
	SplatGeo sg;
	sg.R = r->R;
	sg.rad = r->radius;
	sg.spacing = 0.7 / r->R;
	unsigned pb = ggbk_blocks(np, PBS);
	if (ggbk_zero(r->tile_count, sizeof(int) * (size_t) r->ntiles))
		return -1;
	GGBK_LAUNCH(k_render_count, pb, PBS, 0, np, prev, pos, sg, r->tile, r->ntx, r->tile_count);
	GGBK_LAUNCH(k_render_scan, 1, SCAN_BS, 0, r->ntiles, (const int *) r->tile_count, r->tile_start,
		    r->tile_cursor, r->total);
	long long stats[2] = { 0, 0 };
	if (ggbk_d2h(stats, r->total, sizeof(stats)))
		return -1;
	if (stats[0] > r->cap || stats[0] > INT_MAX / 2) {
		if (stats[0] > INT_MAX / 2)
			return -1;
		long long ncap = stats[0] + stats[0] / 4 + 1024;
		ggbk_free(r->list);
		ggbk_free(r->sort_scratch);
		r->list = (int *) ggbk_malloc(sizeof(int) * (size_t) ncap);
		r->sort_scratch = (int *) ggbk_malloc(sizeof(int) * 2 * (size_t) ncap);
		if (!r->list || !r->sort_scratch)
			return -1;
		r->cap = ncap;
		r->scratch_cap = 2 * ncap;
	}
	GGBK_LAUNCH(k_render_fill, pb, PBS, 0, np, prev, pos, sg, r->tile, r->ntx, (const int *) r->tile_start,
		    r->tile_cursor, r->list, r->cap);
	GGBK_LAUNCH(k_render_draw, (unsigned) r->ntiles, DBS, 0, *r, prev, pos, col, sg, invg,
		    (const long long *) r->total);
	if (ggbk_check("render_deposit"))
		return -1;
	return stats[0];
}
