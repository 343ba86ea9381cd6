// render_k.hip -- deterministic trail splatting (reference src/render.c:87-218).
//
// The reference emits one record per (sample, face, 32-row strip), sorts them
// bin-major / thread-minor and blends each bin in order, so every texel sees its
// splats in (particle index, sample) order.  Here the unit of work is a face TILE of
// TILE x TILE texels held in LDS:
//   1. count  -- each particle adds itself once to every tile that the bounding box of
//                its arc's footprints (per face) overlaps; for short arcs the box comes
//                from the two endpoints (the arc is a straight segment on the face);
//   2. scan   -- exclusive scan of the tile counts (one workgroup);
//   3. fill   -- the same boxes write the particle index into each tile's list;
//   4. draw   -- one workgroup per tile: sort the list (bitonic, LDS), load the tile,
//                then per window of WIN particles (ascending index): each particle
//                computes its samples on the tile once (cached in LDS) and appends a
//                2-byte fragment (window rank, sample) to every texel it touches; each
//                texel then sorts its few fragments and blends them in (particle,
//                sample) order -- the reference's per-texel order, with no
//                cross-particle synchronisation.  A window in which a texel collects
//                more than CAPT fragments is retried with a quarter of the particles;
//                a single particle with more than SMAX samples on the tile is blended
//                in order by one thread.
// The texel arithmetic is the reference's (render.c:113-154) in float with no FMA
// contraction, so texels match the CPU path bit for bit for identical positions.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "dev.h"
#include "ggbk.h"

#define PBS 256          // particle kernels
#define DBS 256          // draw workgroup
#define TILE_MAX 32
#define SORT_CAP 1024    // LDS sort capacity per tile
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

// Per-face union of the footprint boxes of a particle's samples (conservative).
struct FaceBoxes {
	int xlo[6], xhi[6], ylo[6], yhi[6];
};

__device__ __forceinline__ void box_add(FaceBoxes &fb, int f, int xlo, int xhi, int ylo, int yhi)
{
	fb.xlo[f] = min(fb.xlo[f], xlo);
	fb.xhi[f] = max(fb.xhi[f], xhi);
	fb.ylo[f] = min(fb.ylo[f], ylo);
	fb.yhi[f] = max(fb.yhi[f], yhi);
}

// exact per-sample boxes of face f
__device__ __forceinline__ void face_box_samples(dv3 a, dv3 b, int ns, int f, const SplatGeo &sg, FaceBoxes &fb)
{
	for (int k = 1; k <= ns; k++) {
		dv3 d = arc_point(a, b, k, ns);
		if (!(face_dot(f, d) >= 0.5))
			continue;
		float px, py;
		face_pixel(f, d, sg.R, px, py);
		int xlo, xhi, ylo, yhi;
		if (footprint(px, py, sg.rad, sg.R, xlo, xhi, ylo, yhi))
			box_add(fb, f, xlo, xhi, ylo, yhi);
	}
}

// Short arcs: every sample d(t) = normalize((1-t) a + t b) has face_dot between
// min(da, db) and max(da, db) / cos(angle / 2), so with both endpoints at dc >= 0.45 the
// arc projects onto face f as the straight segment between the endpoint pixels (great
// circles are gnomonic lines), and with both below 0.45 no sample reaches dc = 0.5.
// Otherwise the samples are walked.  The box is widened by 0.01 texel for rounding.
#define SHORT_ARC 0.05
__device__ __forceinline__ void arc_boxes(dv3 a, dv3 b, const SplatGeo &sg, FaceBoxes &fb)
{
	for (int f = 0; f < 6; f++) {
		fb.xlo[f] = INT_MAX;
		fb.ylo[f] = INT_MAX;
		fb.xhi[f] = INT_MIN;
		fb.yhi[f] = INT_MIN;
	}
	double ang = vangle(a, b);
	int ns = 0;
	bool shortarc = ang < SHORT_ARC;
	for (int f = 0; f < 6; f++) {
		double da = face_dot(f, a), db = face_dot(f, b);
		if (shortarc && da < 0.45 && db < 0.45)
			continue;
		if (shortarc && da >= 0.45 && db >= 0.45) {
			double dc, u, v;
			gnomonic(f, a, dc, u, v);
			double pxa = (u + 1.0) * 0.5 * sg.R - 0.5, pya = (1.0 - v) * 0.5 * sg.R - 0.5;
			gnomonic(f, b, dc, u, v);
			double pxb = (u + 1.0) * 0.5 * sg.R - 0.5, pyb = (1.0 - v) * 0.5 * sg.R - 0.5;
			double x0 = fmin(pxa, pxb) - 0.01 - sg.rad, x1 = fmax(pxa, pxb) + 0.01 + sg.rad;
			double y0 = fmin(pya, pyb) - 0.01 - sg.rad, y1 = fmax(pya, pyb) + 0.01 + sg.rad;
			int xlo = max((int) ceil(x0), 0), xhi = min((int) floor(x1), sg.R - 1);
			int ylo = max((int) ceil(y0), 0), yhi = min((int) floor(y1), sg.R - 1);
			if (xlo <= xhi && ylo <= yhi)
				box_add(fb, f, xlo, xhi, ylo, yhi);
			continue;
		}
		if (ns == 0) {
			ns = (int) ceil(ang / sg.spacing);
			ns = ns < 1 ? 1 : (ns > 4096 ? 4096 : ns);
		}
		face_box_samples(a, b, ns, f, sg, fb);
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
		for (int k = 2; k < 16; k++)
			stats[k] = 0;
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

// Footprint of one splat clipped to the tile (false if it misses the tile).
__device__ __forceinline__ bool tile_clip(const TileRect &tr, float px, float py, float rad, int R, int &xlo, int &xhi,
					  int &ylo, int &yhi)
{
	if (!footprint(px, py, rad, R, xlo, xhi, ylo, yhi))
		return false;
	xlo = max(xlo, tr.x0);
	xhi = min(xhi, tr.x0 + tr.w - 1);
	ylo = max(ylo, tr.y0);
	yhi = min(yhi, tr.y0 + tr.h - 1);
	return xlo <= xhi && ylo <= yhi;
}

#define WIN 128        // particles per window (one thread each)
#define SMAX 12        // samples per particle cached in LDS (k fits 4 bits)
#define CAPT 12        // fragments one texel can hold per window
#define NTEX (TILE_MAX * TILE_MAX)

// Texels of one splat on the tile with positive tent weight (render.c:129-144 tests).
// fn(local texel index) is called for each, in y-then-x order.
template <typename Fn>
__device__ __forceinline__ void splat_texels(const TileRect &tr, float px, float py, float rad, float invrad, int R,
					     Fn fn)
{
	int xlo, xhi, ylo, yhi;
	if (!tile_clip(tr, px, py, rad, R, xlo, xhi, ylo, yhi))
		return;
	for (int y = ylo; y <= yhi; y++) {
		float wy = 1.0f - fabsf((float) y - py) * invrad;
		if (wy <= 0)
			continue;
		for (int x = xlo; x <= xhi; x++) {
			float wx = 1.0f - fabsf((float) x - px) * invrad;
			if (wx <= 0)
				continue;
			fn((y - tr.y0) * TILE_MAX + (x - tr.x0));
		}
	}
}

// One texel's blend of one splat (render.c:145-150).
__device__ __forceinline__ void blend_texel(float4 &dd, int t, const TileRect &tr, float px, float py, float invrad,
					    float w0, float4 c, float invg)
{
	int x = tr.x0 + t % TILE_MAX, y = tr.y0 + t / TILE_MAX;
	float wy = 1.0f - fabsf((float) y - py) * invrad;
	float wx = 1.0f - fabsf((float) x - px) * invrad;
	float w = w0 * wx * wy;
	float keep = 1.0f - w, add = w * invg;
	dd.x = dd.x * keep + add * c.x;
	dd.y = dd.y * keep + add * c.y;
	dd.z = dd.z * keep + add * c.z;
	dd.w = dd.w * keep + add;
}

__global__ __launch_bounds__(DBS) void k_render_draw(ggbk_render r, const double *prev, const double *pos,
						     const float *col, SplatGeo sg, float invg, const long long *stats)
{
	__shared__ float4 tex[NTEX];
	__shared__ int sbuf[SORT_CAP];
	__shared__ float2 samp[SMAX][WIN];
	__shared__ float4 pcol[WIN];
	__shared__ int tcnt[NTEX];
	__shared__ unsigned short frag[CAPT][NTEX];
	__shared__ int serial;
	__shared__ unsigned tmask[NTEX / 32];
	if (stats[0] > r.cap)
		return; // list overflowed: the host grows it before this launch
	long long c0 = clock64();
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
	const float rad = sg.rad, invrad = 1.0f / rad;

	// sorted particle list: in LDS when it fits, else in this tile's scratch region
	int P = 1;
	while (P < L)
		P <<= 1;
	int *sorted = P <= SORT_CAP ? sbuf : r.sort_scratch + 2 * (size_t) r.tile_start[t];
	for (int i = threadIdx.x; i < P; i += DBS)
		sorted[i] = i < L ? lst[i] : INT_MAX;
	__syncthreads();
	bitonic_sort(sorted, P);

	long long c1 = clock64();
	float4 *gbuf = (float4 *) r.buf;
	for (int i = threadIdx.x; i < NTEX; i += DBS) {
		int ly = i / TILE_MAX, lx = i % TILE_MAX;
		if (ly < tr.h && lx < tr.w)
			tex[i] = gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx];
	}
	for (int i = threadIdx.x; i < NTEX / 32; i += DBS)
		tmask[i] = 0;
	long long csamp = 0, cblend = 0;
	int windows = 0, serial_windows = 0;
	// windows of W particles; a window whose fragments overflow is retried with W / 4
	// particles (the tile is only written by a window that fits), down to one particle,
	// which the serial path always handles
	int W = WIN;
	for (int base = 0; base < L;) {
		long long ca = clock64();
		windows++;
		for (int i = threadIdx.x; i < NTEX; i += DBS)
			tcnt[i] = 0;
		if (threadIdx.x == 0)
			serial = 0;
		__syncthreads();
		// samples of this window's particles on the tile (one thread per particle); every
		// (particle rank, sample, texel) fragment goes to its texel's slot list
		const int rank = threadIdx.x;
		if (rank < W && base + rank < L) {
			int pid = sorted[base + rank];
			dv3 a = ld3(prev + 3 * (size_t) pid), b = ld3(pos + 3 * (size_t) pid);
			int ns = arc_samples(a, b, sg.spacing);
			int nsamp = 0;
			bool over = false;
			for (int k = 1; k <= ns; k++) {
				float px, py;
				if (!face_pixel(tr.f, arc_point(a, b, k, ns), sg.R, px, py))
					continue;
				int xlo, xhi, ylo, yhi;
				if (!tile_clip(tr, px, py, rad, sg.R, xlo, xhi, ylo, yhi))
					continue;
				if (nsamp >= SMAX) {
					over = true;
					break;
				}
				samp[nsamp][rank] = make_float2(px, py);
				unsigned short key = (unsigned short) ((rank << 4) | nsamp);
				splat_texels(tr, px, py, rad, invrad, sg.R, [&](int tx) {
					atomicOr(&tmask[tx >> 5], 1u << (tx & 31));
					int slot = atomicAdd(&tcnt[tx], 1);
					if (slot < CAPT)
						frag[slot][tx] = key;
					else
						over = true;
				});
				nsamp++;
			}
			const float *cp = col + 4 * (size_t) pid;
			pcol[rank] = make_float4(cp[0], cp[1], cp[2], cp[3]);
			if (over)
				serial = 1;
		}
		__syncthreads();
		long long cb = clock64();
		csamp += cb - ca;
		if (!serial) {
			// each texel applies its fragments in (particle rank, sample) order
			for (int tx = threadIdx.x; tx < NTEX; tx += DBS) {
				int n = tcnt[tx];
				if (n == 0)
					continue;
				unsigned short keys[CAPT];
#pragma unroll
				for (int a = 0; a < CAPT; a++)
					keys[a] = a < n ? frag[a][tx] : (unsigned short) 0xffff;
#pragma unroll
				for (int a = 1; a < CAPT; a++) {
					unsigned short key = keys[a];
#pragma unroll
					for (int bpos = a; bpos > 0; bpos--) {
						if (keys[bpos - 1] > key) {
							keys[bpos] = keys[bpos - 1];
							keys[bpos - 1] = key;
						}
					}
				}
				float4 dd = tex[tx];
				for (int a = 0; a < n; a++) {
					int pr = keys[a] >> 4, k = keys[a] & 15;
					float2 q = samp[k][pr];
					float4 c = pcol[pr];
					blend_texel(dd, tx, tr, q.x, q.y, invrad, r.alpha * c.w, c, invg);
				}
				tex[tx] = dd;
			}
			base += W;
			W = WIN;
		} else if (W > 1) {
			W = W >= 4 ? W / 4 : 1;
		} else {
			// one particle with more than SMAX samples on the tile, or more than CAPT
			// samples on one texel: one thread blends it in order
			serial_windows++;
			if (threadIdx.x == 0) {
				for (int rr = 0; rr < 1 && base + rr < L; rr++) {
					int pid = sorted[base + rr];
					dv3 a = ld3(prev + 3 * (size_t) pid), b = ld3(pos + 3 * (size_t) pid);
					int ns = arc_samples(a, b, sg.spacing);
					const float *cp = col + 4 * (size_t) pid;
					float4 c = make_float4(cp[0], cp[1], cp[2], cp[3]);
					for (int k = 1; k <= ns; k++) {
						float px, py;
						if (!face_pixel(tr.f, arc_point(a, b, k, ns), sg.R, px, py))
							continue;
						splat_texels(tr, px, py, rad, invrad, sg.R, [&](int tx) {
							tmask[tx >> 5] |= 1u << (tx & 31);
							blend_texel(tex[tx], tx, tr, px, py, invrad, r.alpha * c.w, c, invg);
						});
					}
				}
			}
			base += 1;
			W = WIN;
		}
		__syncthreads();
		cblend += clock64() - cb;
	}

	// write back the texels a splat touched
	for (int i = threadIdx.x; i < NTEX; i += DBS) {
		int ly = i / TILE_MAX, lx = i % TILE_MAX;
		if ((tmask[i >> 5] >> (i & 31)) & 1u)
			gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx] = tex[i];
	}
	// per-launch diagnostics: [2] windows, [3] serial windows, [4] sort cycles,
	// [5] sample cycles, [6] blend cycles, [7] total cycles (sums over tiles),
	// [8] touched texels
	if (threadIdx.x < NTEX / 32) {
		int c = __popc(tmask[threadIdx.x]);
		if (c)
			atomicAdd((unsigned long long *) &((long long *) stats)[8], (unsigned long long) c);
	}
	if (threadIdx.x == 0) {
		long long *st = (long long *) stats;
		atomicAdd((unsigned long long *) &st[2], (unsigned long long) windows);
		atomicAdd((unsigned long long *) &st[3], (unsigned long long) serial_windows);
		atomicAdd((unsigned long long *) &st[4], (unsigned long long) (c1 - c0));
		atomicAdd((unsigned long long *) &st[5], (unsigned long long) csamp);
		atomicAdd((unsigned long long *) &st[6], (unsigned long long) cblend);
		atomicAdd((unsigned long long *) &st[7], (unsigned long long) (clock64() - c0));
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
extern "C" int ggbk_render_stats(const ggbk_render *r, long long *out16)
{
	return ggbk_d2h(out16, r->total, 16 * sizeof(long long));
}

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
	int pr_bin = ggbk_probe_begin(GGBK_PROBE_RENDER_BIN);
	GGBK_LAUNCH(k_render_count, pb, PBS, 0, np, prev, pos, sg, r->tile, r->ntx, r->tile_count);
	GGBK_LAUNCH(k_render_scan, 1, SCAN_BS, 0, r->ntiles, (const int *) r->tile_count, r->tile_start,
		    r->tile_cursor, r->total);
	ggbk_probe_end(GGBK_PROBE_RENDER_BIN);
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
	if (pr_bin)
		ggbk_probe_begin(GGBK_PROBE_RENDER_BIN);
	GGBK_LAUNCH(k_render_fill, pb, PBS, 0, np, prev, pos, sg, r->tile, r->ntx, (const int *) r->tile_start,
		    r->tile_cursor, r->list, r->cap);
	ggbk_probe_end(GGBK_PROBE_RENDER_BIN);
	ggbk_probe_begin(GGBK_PROBE_RENDER_DRAW);
	GGBK_LAUNCH(k_render_draw, (unsigned) r->ntiles, DBS, 0, *r, prev, pos, col, sg, invg,
		    (const long long *) r->total);
	ggbk_probe_end(GGBK_PROBE_RENDER_DRAW);
	if (ggbk_check("render_deposit"))
		return -1;
	return stats[0];
}
