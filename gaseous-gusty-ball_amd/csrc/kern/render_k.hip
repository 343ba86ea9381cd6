// render_k.hip -- deterministic trail splatting (reference src/render.c:87-218).
//
// The reference emits one record per (sample, face, 32-row strip), sorts them
// bin-major / thread-minor and blends each bin in order, so every texel sees its
// splats in (particle index, sample) order.  Here the unit of work is a face TILE of
// TILE x TILE texels, and the particle arrays are kept in tile order (storage slot s
// holds particle pid[s]), restored every few deposits:
//   1. count  -- (reordering deposits) each particle adds itself to its home tile (the
//                tile of the texel nearest its current position); atomics are
//                aggregated per wave, one per distinct tile, since neighbouring slots
//                hold neighbouring particles; then exclusive scans of the counts;
//   2. fill   -- every tile that the bounding box of the particle's arc footprints (per
//                face) overlaps gets the particle's slot in its fixed-size bucket, and
//                every tile is counted exactly.  For short arcs the box comes from the
//                two endpoints (the arc is a straight segment on the face).  When
//                reordering, each particle is also copied to its home tile's range of
//                the next storage order, which the caller adopts after the draw;
//   3. draw   -- one workgroup per tile, its texels in registers: per window of
//                particles, every (particle, sample) pixel on the tile with one thread
//                per sample, each on-tile sample marked in the texels it touches; each
//                texel blends its samples in (particle index, sample) order -- the
//                reference's per-texel order.  A tile with more
//                than a bucket of particles takes the packed-list path (exact counts,
//                scanned starts, the fill again), is sorted by particle index and cut
//                into windows.
// The texel arithmetic is the reference's (render.c:113-154) in float with no FMA
// contraction, so texels match the CPU path bit for bit for identical positions.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <type_traits>

#include "dev.h"
#include "ggbk.h"

#define PBS 256           // particle kernels
#define TILE 16
#define NTEX (TILE * TILE)
#define WIN 128           // particles per window = draw workgroup size
#define DBS WIN
#define BUCKET (2 * WIN)  // slots per tile bucket: two draw windows
#define SCAN_BS 1024

struct SplatGeo {
	int R, ntx;
	float rad;
	double spacing;
};

__device__ __forceinline__ int arc_samples(double ang, double spacing)
{
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
	if (!(face_dot(f, d) >= 0.5))
		return false;
	double dc, u, v;
	gnomonic(f, d, dc, u, v);
	px = (float) ((u + 1.0) * 0.5 * R - 0.5);
	py = (float) ((1.0 - v) * 0.5 * R - 0.5);
	return true;
}

// face_pixel for a face fixed at compile time (the draw's tile face, chosen once per
// range): the same arithmetic, with the component selection resolved by the compiler.
template <int F>
__device__ __forceinline__ bool face_pixel_c(dv3 d, int R, float &px, float &py)
{
	double dc, cr, cu;
	if constexpr (F == 0) {
		dc = d.z; cr = d.x; cu = d.y;
	} else if constexpr (F == 1) {
		dc = d.x; cr = -d.z; cu = d.y;
	} else if constexpr (F == 2) {
		dc = -d.z; cr = -d.x; cu = d.y;
	} else if constexpr (F == 3) {
		dc = -d.x; cr = d.z; cu = d.y;
	} else if constexpr (F == 4) {
		dc = d.y; cr = d.x; cu = -d.z;
	} else {
		dc = -d.y; cr = d.x; cu = d.z;
	}
	if (!(dc >= 0.5))
		return false;
	double u = cr / dc, v = cu / dc;
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

// Pixel coordinates of direction d on face f in single precision, through one hardware
// reciprocal (about 1 ulp): for the binning only, whose boxes carry a margin far above
// this error (at dc >= 0.45 and R <= 4096, under 0.002 texel), and for the home tile,
// which only orders storage.  count and fill evaluate it identically.
__device__ __forceinline__ void pixel_fast(int f, dv3 d, int R, float &px, float &py)
{
	double dc, cr, cu;
	switch (f) {
	case 0: dc = d.z; cr = d.x; cu = d.y; break;
	case 1: dc = d.x; cr = -d.z; cu = d.y; break;
	case 2: dc = -d.z; cr = -d.x; cu = d.y; break;
	case 3: dc = -d.x; cr = d.z; cu = d.y; break;
	case 4: dc = d.y; cr = d.x; cu = -d.z; break;
	default: dc = -d.y; cr = d.x; cu = d.z; break;
	}
	float inv = __builtin_amdgcn_rcpf((float) dc);
	float u = (float) cr * inv, v = (float) cu * inv;
	float hR = 0.5f * (float) R;
	px = (u + 1.0f) * hR - 0.5f;
	py = (1.0f - v) * hR - 0.5f;
}

// Tile of the texel nearest the particle's current position, on the face it lies on.
__device__ __forceinline__ int home_tile(dv3 p, int R, int ntx)
{
	int f = face_of(p);
	float px, py;
	pixel_fast(f, p, R, px, py);
	int ix = (int) fminf(fmaxf(floorf(px + 0.5f), 0.0f), (float) (R - 1));
	int iy = (int) fminf(fmaxf(floorf(py + 0.5f), 0.0f), (float) (R - 1));
	return (f * ntx + iy / TILE) * ntx + ix / TILE;
}

// Union over the arc's samples of their footprint boxes on face f (conservative);
// false if no sample reaches the face.  `ns` is the sample count, computed on first use
// (0 = not yet).
//
// Short arcs: every sample d(t) = normalize((1-t) a + t b) has face_dot between
// min(da, db) and max(da, db) / cos(angle / 2), so with both endpoints at dc >= 0.45 the
// arc projects onto face f as the straight segment between the endpoint pixels (great
// circles are gnomonic lines), and with both below 0.45 no sample reaches dc = 0.5
// (both hold for any angle below 0.9 rad; "short" is a chord under 2 sin(0.025), an
// angle under 0.05).  Otherwise the samples are walked.  The box is widened by 0.01
// texel for rounding.
#define SHORT_CHORD2 (4.0 * 0.024997395914712332 * 0.024997395914712332)
__device__ __forceinline__ bool face_box(dv3 a, dv3 b, bool shortarc, int &ns, int f, const SplatGeo &sg, int &xlo,
					 int &xhi, int &ylo, int &yhi)
{
	double da = face_dot(f, a), db = face_dot(f, b);
	if (shortarc && da < 0.45 && db < 0.45)
		return false;
	if (shortarc && da >= 0.45 && db >= 0.45) {
		float pxa, pya, pxb, pyb;
		pixel_fast(f, a, sg.R, pxa, pya);
		pixel_fast(f, b, sg.R, pxb, pyb);
		float x0 = fminf(pxa, pxb) - 0.01f - sg.rad, x1 = fmaxf(pxa, pxb) + 0.01f + sg.rad;
		float y0 = fminf(pya, pyb) - 0.01f - sg.rad, y1 = fmaxf(pya, pyb) + 0.01f + sg.rad;
		// clamp before converting (dc >= 0.45 keeps |u|, |v| below 2.1)
		xlo = (int) fmaxf(ceilf(x0), 0.0f);
		xhi = (int) fminf(floorf(x1), (float) (sg.R - 1));
		ylo = (int) fmaxf(ceilf(y0), 0.0f);
		yhi = (int) fminf(floorf(y1), (float) (sg.R - 1));
		return xlo <= xhi && ylo <= yhi;
	}
	if (ns == 0)
		ns = arc_samples(vangle(a, b), sg.spacing);
	xlo = ylo = INT_MAX;
	xhi = yhi = INT_MIN;
	for (int k = 1; k <= ns; k++) {
		float px, py;
		if (!face_pixel(f, arc_point(a, b, k, ns), sg.R, px, py))
			continue;
		int x0, x1, y0, y1;
		if (footprint(px, py, sg.rad, sg.R, x0, x1, y0, y1)) {
			xlo = min(xlo, x0);
			xhi = max(xhi, x1);
			ylo = min(ylo, y0);
			yhi = max(yhi, y1);
		}
	}
	return xlo <= xhi;
}

// Adds one per participating lane to cnt[key] with one atomic per distinct key in the
// wave and returns each participating lane's slot (previous count + rank among the lanes
// with the same key, in lane order); -1 for lanes not participating.  Called by every
// lane of the wave.  The distinct keys are found with ballots first; their leaders then
// issue all the atomics in one instruction, so the wave waits for one round trip.
__device__ __forceinline__ int wave_atomic_inc(int *cnt, int key, bool on)
{
	unsigned long long todo = __ballot(on);
	int lane = __lane_id();
	int leader_of = lane, rank = 0, add = 0;
	while (todo) {
		int leader = __ffsll((long long) todo) - 1;
		int lk = __builtin_amdgcn_readlane(key, leader);
		unsigned long long m = __ballot(on && key == lk) & todo;
		if (lane == leader)
			add = __popcll(m);
		if ((m >> lane) & 1ull) {
			leader_of = leader;
			rank = __popcll(m & ((1ull << lane) - 1ull));
		}
		todo &= ~m;
	}
	int base = 0;
	if (add > 0)
		base = atomicAdd(&cnt[key], add);
	base = __shfl(base, leader_of);
	return on ? base + rank : -1;
}

// The same without the slots: the atomics return nothing, so the wave does not wait for
// each of them in turn.
__device__ __forceinline__ void wave_atomic_count(int *cnt, int key, bool on)
{
	unsigned long long todo = __ballot(on);
	int lane = __lane_id();
	while (todo) {
		int leader = __ffsll((long long) todo) - 1;
		int lk = __builtin_amdgcn_readlane(key, leader);
		unsigned long long m = __ballot(on && key == lk) & todo;
		if (lane == leader)
			atomicAdd(&cnt[lk], __popcll(m));
		todo &= ~m;
	}
}

// Visits every tile of every face box of particle i's arc: fn(has, tile) is called by
// all lanes of the wave together, once per tile index step, with has false for lanes
// that have no tile at this step.
template <typename Fn>
__device__ __forceinline__ void for_arc_tiles(bool on, dv3 a, dv3 b, const SplatGeo &sg, Fn fn)
{
	dv3 dab = vsub(a, b);
	bool shortarc = vdot(dab, dab) < SHORT_CHORD2;
	int ns = 0;
	// faces this lane's arc can reach (a short arc only those with an endpoint at
	// dc >= 0.45; a long one any), united over the wave: the loop visits only those
	unsigned fmask = 0;
	if (on) {
		for (int f = 0; f < 6; f++)
			if (!shortarc || face_dot(f, a) >= 0.45 || face_dot(f, b) >= 0.45)
				fmask |= 1u << f;
	}
	unsigned wmask = 0;
	for (int f = 0; f < 6; f++)
		if (__ballot((fmask >> f) & 1u))
			wmask |= 1u << f;
	fmask = wmask;
	for (int f = 0; f < 6; f++) {
		if (!((fmask >> f) & 1u))
			continue;
		int xlo = 0, xhi = -1, ylo = 0, yhi = -1;
		bool any = on && face_box(a, b, shortarc, ns, f, sg, xlo, xhi, ylo, yhi);
		int tx0 = xlo / TILE, ty0 = ylo / TILE;
		int w = any ? xhi / TILE - tx0 + 1 : 0;
		int nt = any ? w * (yhi / TILE - ty0 + 1) : 0;
		// tile (tx0 + ix, ty0 + iy) at step s = iy w + ix, stepped without a division
		int ix = 0, iy = 0;
		for (int s = 0; __any(s < nt); s++) {
			bool has = s < nt;
			int t = has ? (f * sg.ntx + ty0 + iy) * sg.ntx + tx0 + ix : 0;
			fn(has, t);
			if (++ix == w) {
				ix = 0;
				iy++;
			}
		}
	}
}

// Particles per home tile (the next storage order).  The tile lists are counted by the
// fill itself, into fixed-size buckets.
__global__ __launch_bounds__(PBS) void k_render_count(ggbk_parts ps, SplatGeo sg, int *hcounts)
{
	int i = xblock() * PBS + threadIdx.x;
	bool on = i < ps.n;
	dv3 b = d3(0, 0, 1);
	if (on)
		b = ld3(ps.pos + 3 * (size_t) i);
	int home = on ? home_tile(b, sg.R, sg.ntx) : 0;
	wave_atomic_count(hcounts, home, on);
}

// Writes each particle's (particle, tile) entries, and its copy in the next storage order.
// With bucket != NULL the entries go to fixed-size tile buckets and cursor ends as the
// exact count of every tile (entries past a full bucket are counted, not stored);
// otherwise to the list at the scanned starts.  cursor starts at zero (the previous
// deposit's draw cleared it, or the scan did).  The bucket fill also clears the deposit's
// statistics, which the bucket statistics after it accumulate into.
__global__ __launch_bounds__(PBS) void k_render_fill(ggbk_parts ps, ggbk_parts out, SplatGeo sg, const int *starts,
						     int *cursor, const int *hstarts, int *hcursor, int *list,
						     long long cap, int *bucket, unsigned short *nsamp, long long *stats)
{
	if (bucket && blockIdx.x == 0 && threadIdx.x < 16)
		stats[threadIdx.x] = 0;
	int i = xblock() * PBS + threadIdx.x;
	bool on = i < ps.n;
	dv3 a = d3(0, 0, 1), b = d3(0, 0, 1);
	if (on) {
		a = ld3(ps.prev + 3 * (size_t) i);
		b = ld3(ps.pos + 3 * (size_t) i);
	}
	// the arc's sample count (render.c:179-181), for the draw: it needs the exact angle,
	// whose FP64 atan2 would double the draw kernel's registers
	if (on)
		nsamp[i] = (unsigned short) arc_samples(vangle(a, b), sg.spacing);
	// the particle's place in the next storage order (hcursor NULL: no reordering)
	if (hcursor) {
		int home = on ? home_tile(b, sg.R, sg.ntx) : 0;
		int hr = wave_atomic_inc(hcursor, home, on);
		if (on) {
			size_t o = (size_t) hstarts[home] + hr;
			st3(out.pos + 3 * o, b);
			st3(out.prev + 3 * o, a);
			*(float4 *) (out.col + 4 * o) = *(const float4 *) (ps.col + 4 * (size_t) i);
			out.pid[o] = ps.pid[i];
		}
	}
	for_arc_tiles(on, a, b, sg, [&](bool has, int t) {
		int r = wave_atomic_inc(cursor, t, has);
		if (has) {
			if (bucket) {
				if (r < BUCKET)
					bucket[(size_t) t * BUCKET + r] = i;
			} else {
				long long slot = (long long) starts[t] + r;
				if (slot < cap)
					list[slot] = i;
			}
		}
	});
}

// Inclusive scan over the 64 lanes of a wave with DPP moves (no LDS): Hillis-Steele
// within each row of 16 lanes (row_shr 1, 2, 4, 8; lanes shifted in from outside the row
// add 0), then row 1 and row 3 add the last lane of the row before (row_bcast:15), and
// rows 2 and 3 add lane 31 (row_bcast:31).
__device__ __forceinline__ int wave_incl_scan(int v)
{
	v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
	return v;
}

// Exclusive scans of the tile list counts and the home counts, in two passes over
// blocks of SCAN_BS tiles: local scans plus per-block totals, then each block adds the
// totals of the blocks before it.  stats[0] = total list entries, stats[1] = longest
// tile list; stats[2..15] are cleared.
__global__ __launch_bounds__(SCAN_BS) void k_render_scan_local(int ntiles, const int *counts, int *starts,
							       int *cursor, const int *hcounts, int *hstarts,
							       int *hcursor, long long *part)
{
	__shared__ int wsum[SCAN_BS / 64], hwsum[SCAN_BS / 64], wmax[SCAN_BS / 64];
	int t = blockIdx.x * SCAN_BS + threadIdx.x;
	int c = t < ntiles ? counts[t] : 0;
	int hc = t < ntiles ? hcounts[t] : 0;
	int ic = wave_incl_scan(c), ihc = wave_incl_scan(hc);
	int m = c;
	for (int off = 32; off > 0; off >>= 1)
		m = max(m, __shfl_xor(m, off));
	int w = threadIdx.x / 64, lane = __lane_id();
	if (lane == 63) {
		wsum[w] = ic;
		hwsum[w] = ihc;
		wmax[w] = m;
	}
	__syncthreads();
	int before = 0, hbefore = 0;
	for (int k = 0; k < w; k++) {
		before += wsum[k];
		hbefore += hwsum[k];
	}
	if (t < ntiles) {
		starts[t] = before + ic - c;
		hstarts[t] = hbefore + ihc - hc;
		cursor[t] = 0;
		hcursor[t] = 0;
	}
	if (threadIdx.x == 0) {
		long long s = 0, hs = 0;
		int mm = 0;
		for (int k = 0; k < SCAN_BS / 64; k++) {
			s += wsum[k];
			hs += hwsum[k];
			mm = max(mm, wmax[k]);
		}
		part[3 * (size_t) blockIdx.x] = s;
		part[3 * (size_t) blockIdx.x + 1] = hs;
		part[3 * (size_t) blockIdx.x + 2] = mm;
	}
}

__global__ __launch_bounds__(SCAN_BS) void k_render_scan_fix(int ntiles, int *starts, int *hstarts,
							     const long long *part, long long *stats)
{
	__shared__ long long red[3][SCAN_BS / 64];
	int b = blockIdx.x, nb = gridDim.x;
	bool last = b == nb - 1;
	long long s = 0, hs = 0, m = 0;
	for (int j = threadIdx.x; j < (last ? nb : b); j += SCAN_BS) {
		if (j < b) {
			s += part[3 * (size_t) j];
			hs += part[3 * (size_t) j + 1];
		}
		m = max(m, part[3 * (size_t) j + 2]);
	}
	for (int off = 32; off > 0; off >>= 1) {
		s += __shfl_xor(s, off);
		hs += __shfl_xor(hs, off);
		m = max(m, __shfl_xor(m, off));
	}
	int w = threadIdx.x / 64;
	if (__lane_id() == 0) {
		red[0][w] = s;
		red[1][w] = hs;
		red[2][w] = m;
	}
	__syncthreads();
	s = hs = m = 0;
	for (int k = 0; k < SCAN_BS / 64; k++) {
		s += red[0][k];
		hs += red[1][k];
		m = max(m, red[2][k]);
	}
	int t = b * SCAN_BS + threadIdx.x;
	if (t < ntiles) {
		starts[t] += (int) s;
		hstarts[t] += (int) hs;
	}
	if (last && threadIdx.x == 0) {
		long long tot = s + part[3 * (size_t) b], htot = hs + part[3 * (size_t) b + 1];
		starts[ntiles] = (int) tot;
		hstarts[ntiles] = (int) htot;
		stats[0] = tot;
		stats[1] = m;
		for (int k = 2; k < 16; k++)
			stats[k] = 0;
	}
}

// stats[0] += total of the bucket counts, stats[1] = max(stats[1], longest) (integer
// atomics: order-free).
__global__ __launch_bounds__(SCAN_BS) void k_render_bucket_stats(int ntiles, const int *counts, long long *stats)
{
	__shared__ long long ws[SCAN_BS / 64];
	__shared__ int wm[SCAN_BS / 64];
	int t = blockIdx.x * SCAN_BS + threadIdx.x;
	long long c = t < ntiles ? counts[t] : 0;
	int m = (int) c;
	for (int off = 32; off > 0; off >>= 1) {
		c += __shfl_xor(c, off);
		m = max(m, __shfl_xor(m, off));
	}
	if (__lane_id() == 0) {
		ws[threadIdx.x / 64] = c;
		wm[threadIdx.x / 64] = m;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		long long sum = 0;
		int mm = 0;
		for (int k = 0; k < SCAN_BS / 64; k++) {
			sum += ws[k];
			mm = max(mm, wm[k]);
		}
		atomicAdd((unsigned long long *) &stats[0], (unsigned long long) sum);
		atomicMax(&stats[1], (long long) mm);
	}
}

// bitonic sort of P (a power of two) 64-bit keys, ascending
__device__ __forceinline__ void bitonic_sort64(unsigned long long *a, int P)
{
	for (int k = 2; k <= P; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < P; i += DBS) {
				int ixj = i ^ j;
				if (ixj > i) {
					unsigned long long x = a[i], y = a[ixj];
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

// Texels of one splat on the tile with positive tent weight (render.c:129-144 tests);
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
			fn((y - tr.y0) * TILE + (x - tr.x0));
		}
	}
}

// splat_texels for radius <= 1, from the splat's clipped box [xlo, xhi] x [ylo, yhi]:
// the texels with positive weight (|x - px| * invrad < 1) are at most two consecutive
// columns and two consecutive rows, starting within two of the box's first column and
// row.  The (up to) four are returned at once, on[q] / tx[q], so that their LDS atomics
// can be issued together; the order of a splat's own texels does not matter.
__device__ __forceinline__ void splat_texels_2x2(const TileRect &tr, float px, float py, float invrad, int xlo,
						 int xhi, int ylo, int yhi, bool on[4], int tx[4])
{
	auto pos = [&](int v, float c, int hi) { return v <= hi && 1.0f - fabsf((float) v - c) * invrad > 0; };
	bool x0 = pos(xlo, px, xhi), x1 = pos(xlo + 1, px, xhi), x2 = pos(xlo + 2, px, xhi);
	bool y0 = pos(ylo, py, yhi), y1 = pos(ylo + 1, py, yhi), y2 = pos(ylo + 2, py, yhi);
	int xa = x0 ? xlo : (x1 ? xlo + 1 : xlo + 2), ya = y0 ? ylo : (y1 ? ylo + 1 : ylo + 2);
	bool ca = x0 || x1 || x2, cb = x0 ? x1 : (x1 && x2);
	bool ra = y0 || y1 || y2, rb = y0 ? y1 : (y1 && y2);
	int base = (ya - tr.y0) * TILE + (xa - tr.x0);
	on[0] = ra && ca;
	tx[0] = base;
	on[1] = ra && cb;
	tx[1] = base + 1;
	on[2] = rb && ca;
	tx[2] = base + TILE;
	on[3] = rb && cb;
	tx[3] = base + TILE + 1;
}

// Whether local texel t is one of splat_texels' texels for the splat at (px, py).
__device__ __forceinline__ bool texel_in_splat(const TileRect &tr, int t, float px, float py, float rad, float invrad,
					       int R)
{
	int xlo, xhi, ylo, yhi;
	if (!tile_clip(tr, px, py, rad, R, xlo, xhi, ylo, yhi))
		return false;
	int x = tr.x0 + t % TILE, y = tr.y0 + t / TILE;
	if (x < xlo || x > xhi || y < ylo || y > yhi)
		return false;
	return 1.0f - fabsf((float) y - py) * invrad > 0 && 1.0f - fabsf((float) x - px) * invrad > 0;
}

// One texel's blend of one splat (render.c:145-150).
__device__ __forceinline__ void blend_texel(float4 &dd, int t, const TileRect &tr, float px, float py, float invrad,
					    float w0, float4 c, float invg)
{
	int x = tr.x0 + t % TILE, y = tr.y0 + t / TILE;
	float wy = 1.0f - fabsf((float) y - py) * invrad;
	float wx = 1.0f - fabsf((float) x - px) * invrad;
	float w = w0 * wx * wy;
	float keep = 1.0f - w, add = w * invg;
	dd.x = dd.x * keep + add * c.x;
	dd.y = dd.y * keep + add * c.y;
	dd.z = dd.z * keep + add * c.z;
	dd.w = dd.w * keep + add;
}

// Per window of W particles (W <= WIN):
//   A  one thread per particle: endpoints, sample count ns, index, colour;
//   B  the particle's rank by index within the window (the window is index-sorted when the
//      tile list was, otherwise ranks are counted); everything is stored by rank, and the
//      exclusive prefix of ns in rank order;
//   then per range of consecutive ranks whose samples fit SCAP (ranges are in index order):
//   C  one thread per (particle, sample), its rank read from a per-range sample map: the
//      sample's pixel on this tile's face; a sample on the tile stores its position in
//      the slot of its index in the range and sets the slot's bit in the mask of every
//      texel it touches.  Slots are numbered in (rank, sample) order;
//   D  one thread per texel: blend the samples of its set bits in ascending slot order,
//      which is the (particle index, sample) order.  A texel takes any number of
//      samples of a range: no capacity to overflow, no retries.
// A single particle with more samples than a range holds is blended in sample order by
// every thread for its own texels.
// The sample counts come from the binning (an FP64 atan2 here would double the
// registers) and each thread keeps its two texels in registers.  The LDS arrays (14.7 KB,
// allocated in 1280-byte granules) leave room for 10 workgroups per CU, and the launch
// bound of 5 waves per SIMD holds the registers to 96, which is the same 10.  Ranges of
// 128 slots measured faster over a developed flow than 256 (8 workgroups per CU), 192
// (9) and 64 (more ranges per window).
#define SCAP 128         // on-tile samples per range: 4 mask words per texel
static_assert(SCAP % 64 == 0 && WIN <= 256, "texel masks are whole 64-bit words; ranks fit the sample map");
__global__ __launch_bounds__(DBS, 5) void k_render_draw(ggbk_render r, ggbk_parts ps, SplatGeo sg, float invg)
{
	static_assert(NTEX % DBS == 0, "each thread owns NTEX / DBS texels");
	constexpr int TPT = NTEX / DBS;
	__shared__ unsigned smask[SCAP / 32][NTEX];  // bit sl: range slot sl touches the texel
	__shared__ float2 sxy[SCAP];
	__shared__ double pa[3][WIN], pb[3][WIN];  // by rank
	__shared__ int pst[WIN + 1];
	__shared__ __align__(16) int ppid[WIN];    // window order; padded to 4 with INT_MAX
	__shared__ unsigned short pns[WIN];        // sample count by rank (<= 4096)
	__shared__ unsigned char smap[SCAP];       // rank - r0 of each sample of a range
	__shared__ float4 pcol[WIN];               // by rank
	__shared__ int ntouched;
	long long c0 = clock64();
	int t = xblock();
	const int lid = threadIdx.x;
	int *diag = r.tile_diag + 8 * (size_t) t;
	TileRect tr;
	tr.f = t / (r.ntx * r.ntx);
	int rem = t % (r.ntx * r.ntx);
	tr.y0 = (rem / r.ntx) * TILE;
	tr.x0 = (rem % r.ntx) * TILE;
	tr.w = min(TILE, r.R - tr.x0);
	tr.h = min(TILE, r.R - tr.y0);
	const float rad = sg.rad, invrad = 1.0f / rad;

	// The tile's count, the entry total and longest list, the tile's texels (thread lid
	// owns texels lid + h DBS and keeps them in registers) and its first window of bucket
	// entries (one per thread) are loaded together, unconditionally and in bounds (the
	// buckets exist in both modes): one round trip before the particles' data.  Entries
	// past the count are stale and never used.
	const int *cnt_hi = r.bucket_mode ? r.tile_cursor + t : r.tile_start + t + 1;
	const int *cnt_lo = r.bucket_mode ? r.tile_cursor + t : r.tile_start + t;
	int c_hi = *cnt_hi, c_lo = *cnt_lo;
	const longlong2 tot = *(const longlong2 *) r.total;
	float4 *gbuf = (float4 *) r.buf;
	float4 tex[TPT];
	unsigned touched = 0;
#pragma unroll
	for (int h = 0; h < TPT; h++) {
		int i = lid + h * DBS;
		int ly = i / TILE, lx = i % TILE;
		tex[h] = make_float4(0, 0, 0, 0);
		if (ly < tr.h && lx < tr.w)
			tex[h] = gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx];
	}
	static_assert(BUCKET >= DBS && BUCKET <= 2 * WIN, "a bucket holds one or two windows");
	int bslot = r.bucket[(size_t) t * BUCKET + lid];
	// keeps the entry's load above the early return below (the compiler would sink it
	// there, into a round trip of its own)
	asm volatile("" ::"v"(bslot));
	int L = r.bucket_mode ? c_hi : c_hi - c_lo;
	// A bucket-mode launch is made before the host has seen the longest list: when a list
	// overflowed its bucket it draws nothing, and the packed list is drawn after it.  A
	// packed list that did not fit is not drawn either (the host grows it and draws again).
	bool skip = r.bucket_mode ? tot.y > BUCKET : tot.x > r.cap;
	if (L == 0 || skip) {
		if (lid < 8)
			diag[lid] = 0;
		return;
	}
	const int *lst = r.bucket_mode ? r.bucket + (size_t) t * BUCKET : r.list + c_lo;
	if (lid == 0)
		ntouched = 0;

	// a tile with more than WIN particles is cut into index-ordered windows: sort the
	// (index, slot) pairs -- a bucket's (at most two windows) in LDS, in the texel masks'
	// space, with each thread then keeping its slot of both windows in registers (the
	// masks are cleared per window); a packed list's in its part of the scratch array
	unsigned long long *sorted = nullptr;
	bool is_sorted = L > WIN;
	int wslot0 = 0, wslot1 = 0;
	if (is_sorted) {
		int P = 1;
		while (P < L)
			P <<= 1;
		static_assert(sizeof(smask) >= 8 * BUCKET, "a bucket's keys fit the mask space");
		sorted = r.bucket_mode ? (unsigned long long *) smask : r.sort_scratch + 2 * (size_t) c_lo;
		for (int i = lid; i < P; i += DBS) {
			unsigned long long key = ~0ull;
			if (i < L) {
				int slot = lst[i];
				key = ((unsigned long long) (unsigned) ps.pid[slot] << 32) | (unsigned) slot;
			}
			sorted[i] = key;
		}
		__syncthreads();
		bitonic_sort64(sorted, P);
		if (r.bucket_mode) {
			wslot0 = (int) (unsigned) sorted[lid];
			wslot1 = (int) (unsigned) sorted[WIN + lid];
		}
	}
	__syncthreads();

	long long csamp = 0, cblend = 0, cpre = 0;
	int windows = 0, serial_windows = 0;
	for (int base = 0; base < L; base += WIN) {
		int W = min(WIN, L - base);
		long long ca = clock64();
		// A: the window's particles
		float4 mycol = make_float4(0, 0, 0, 0);
		int myns = 0, mypid = 0;
		dv3 a = d3(0, 0, 0), b = d3(0, 0, 0);
		if (lid < W) {
			int slot = !is_sorted ? (r.bucket_mode ? bslot : lst[lid])
					      : r.bucket_mode ? (base == 0 ? wslot0 : wslot1) : (int) (unsigned) sorted[base + lid];
			a = ld3(ps.prev + 3 * (size_t) slot);
			b = ld3(ps.pos + 3 * (size_t) slot);
			myns = r.nsamp[slot];
			mypid = ps.pid[slot];
			mycol = *(const float4 *) (ps.col + 4 * (size_t) slot);
			ppid[lid] = mypid;
		} else if (lid < ((W + 3) & ~3)) {
			ppid[lid] = INT_MAX;
		}
		// the first range's texel masks
		for (int i = lid; i < SCAP / 32 * NTEX / 4; i += DBS)
			((uint4 *) smask)[i] = make_uint4(0, 0, 0, 0);
		__syncthreads();
		// B: rank by index; store by rank
		int rank = lid;
		if (lid < W) {
			if (!is_sorted) {
				// count the smaller indices, four per (broadcast) LDS read
				rank = 0;
				const int4 *p4 = (const int4 *) ppid;
				for (int q = 0; q < (W + 3) / 4; q++) {
					int4 v = p4[q];
					rank += (v.x < mypid) + (v.y < mypid) + (v.z < mypid) + (v.w < mypid);
				}
			}
			pa[0][rank] = a.x; pa[1][rank] = a.y; pa[2][rank] = a.z;
			pb[0][rank] = b.x; pb[1][rank] = b.y; pb[2][rank] = b.z;
			pcol[rank] = mycol;
			pns[rank] = (unsigned short) myns;
		}
		__syncthreads();
		// exclusive prefix of ns in rank order: each wave scans both halves (ranks 0-63 and
		// 64-127) itself, so no barrier passes a total between the waves
		int lane = __lane_id();
		int n0 = lane < W ? pns[lane] : 0;
		int n1 = 64 + lane < W ? pns[64 + lane] : 0;
		int i0 = wave_incl_scan(n0), i1 = wave_incl_scan(n1);
		int s0 = __builtin_amdgcn_readlane(i0, 63);
		int total = s0 + __builtin_amdgcn_readlane(i1, 63);
		int mine = lid < 64 ? n0 : n1;
		int excl = (lid < 64 ? i0 : s0 + i1) - mine;
		if (lid < W)
			pst[lid] = excl;
		if (lid == 0)
			pst[W] = total;
		// the whole window in one range (the usual case): its sample map is written here,
		// so the range starts without a set-up step of its own
		const bool whole = total <= SCAP;
		if (whole && lid < W)
			for (int k = excl; k < excl + mine; k++)
				smap[k] = (unsigned char) lid;
		__syncthreads();
		csamp += clock64() - ca;
		cpre += clock64() - ca;

		bool prepared = whole;
		for (int r0 = 0; r0 < W;) {
			long long cc = clock64();
			windows++;
			// the longest range from r0 whose samples fit SCAP (at least one particle);
			// usually the rest of the window
			int lo = r0 + 1, hi = W;
			if (pst[W] - pst[r0] <= SCAP)
				lo = W;
			while (lo < hi) {
				int mid = (lo + hi + 1) >> 1;
				if (pst[mid] - pst[r0] <= SCAP)
					lo = mid;
				else
					hi = mid - 1;
			}
			const int r1 = lo;
			const int j0 = pst[r0], j1 = pst[r1];
			// a single particle with more samples than a range holds goes to the in-order
			// fallback
			const bool big = j1 - j0 > SCAP;
			if (!(prepared && r0 == 0 && r1 == W)) {
				for (int i = lid; i < SCAP / 32 * NTEX / 4; i += DBS)
					((uint4 *) smask)[i] = make_uint4(0, 0, 0, 0);
				// the rank of every sample of the range (r1 - r0 <= WIN = DBS)
				if (!big && r0 + lid < r1) {
					int p = r0 + lid;
					for (int k = pst[p] - j0; k < pst[p + 1] - j0; k++)
						smap[k] = (unsigned char) lid;
				}
				__syncthreads();
			}
			prepared = false;
			if (!big) {
				// C: one thread per (particle, sample) of ranks [r0, r1), specialised for
				// the tile's face; an on-tile sample stores its pixel in its slot and sets
				// the slot's bit in every texel it touches
				auto sample_pass = [&](auto face) {
					constexpr int F = decltype(face)::value;
					for (int j = j0 + lid; j < j1; j += DBS) {
						int p = r0 + smap[j - j0], k = j - pst[p] + 1;
						dv3 pa3 = d3(pa[0][p], pa[1][p], pa[2][p]), pb3 = d3(pb[0][p], pb[1][p], pb[2][p]);
						float px, py;
						if (!face_pixel_c<F>(arc_point(pa3, pb3, k, pns[p]), sg.R, px, py))
							continue;
						int xlo, xhi, ylo, yhi;
						if (!tile_clip(tr, px, py, rad, sg.R, xlo, xhi, ylo, yhi))
							continue;
						const int sl = j - j0; // the range's samples fit SCAP slots
						sxy[sl] = make_float2(px, py);
						unsigned *mrow = smask[sl >> 5];
						const unsigned bit = 1u << (sl & 31);
						if (rad <= 1.0f) {
							bool on[4];
							int tx[4];
							splat_texels_2x2(tr, px, py, invrad, xlo, xhi, ylo, yhi, on, tx);
#pragma unroll
							for (int q = 0; q < 4; q++)
								if (on[q])
									atomicOr(&mrow[tx[q]], bit);
						} else {
							splat_texels(tr, px, py, rad, invrad, sg.R,
								     [&](int tx) { atomicOr(&mrow[tx], bit); });
						}
					}
				};
				switch (tr.f) {
				case 0: sample_pass(std::integral_constant<int, 0>()); break;
				case 1: sample_pass(std::integral_constant<int, 1>()); break;
				case 2: sample_pass(std::integral_constant<int, 2>()); break;
				case 3: sample_pass(std::integral_constant<int, 3>()); break;
				case 4: sample_pass(std::integral_constant<int, 4>()); break;
				default: sample_pass(std::integral_constant<int, 5>()); break;
				}
				__syncthreads();
				long long cb = clock64();
				csamp += cb - cc;
				// D: each texel blends the samples whose bits it holds, in slot order --
				// the (particle index, sample) order
#pragma unroll
				for (int h = 0; h < TPT; h++) {
					const int tx = lid + h * DBS;
					unsigned m[SCAP / 32], any = 0;
#pragma unroll
					for (int w = 0; w < SCAP / 32; w++) {
						m[w] = smask[w][tx];
						any |= m[w];
					}
					if (!any)
						continue;
					touched |= 1u << h;
					float4 dd = tex[h];
					// 64-bit words: fewer loops whose lanes diverge over the words
#pragma unroll
					for (int w = 0; w < SCAP / 64; w++) {
						unsigned long long bits =
							((unsigned long long) m[2 * w + 1] << 32) | m[2 * w];
						while (bits) {
							const int sl = w * 64 + __ffsll((long long) bits) - 1;
							bits &= bits - 1;
							float2 s2 = sxy[sl];
							float4 c = pcol[r0 + smap[sl]];
							blend_texel(dd, tx, tr, s2.x, s2.y, invrad, r.alpha * c.w, c, invg);
						}
					}
					tex[h] = dd;
				}
				cblend += clock64() - cb;
			} else {
				// one particle with more samples than a range holds: every thread walks
				// them in order and blends those that cover its own texels
				serial_windows++;
				dv3 pa3 = d3(pa[0][r0], pa[1][r0], pa[2][r0]);
				dv3 pb3 = d3(pb[0][r0], pb[1][r0], pb[2][r0]);
				int ns = pns[r0];
				float4 c = pcol[r0];
				for (int k = 1; k <= ns; k++) {
					float px, py;
					if (!face_pixel(tr.f, arc_point(pa3, pb3, k, ns), sg.R, px, py))
						continue;
#pragma unroll
					for (int h = 0; h < TPT; h++) {
						int tx = lid + h * DBS;
						if (texel_in_splat(tr, tx, px, py, rad, invrad, sg.R)) {
							touched |= 1u << h;
							blend_texel(tex[h], tx, tr, px, py, invrad, r.alpha * c.w, c, invg);
						}
					}
				}
			}
			__syncthreads();
			r0 = r1;
		}
	}

	// write back the texels a splat touched (NTEX is a multiple of DBS: whole waves)
	int cnt = 0;
#pragma unroll
	for (int h = 0; h < TPT; h++) {
		int i = lid + h * DBS;
		int ly = i / TILE, lx = i % TILE;
		bool w = (touched >> h) & 1u;
		if (w)
			gbuf[((size_t) tr.f * r.R + tr.y0 + ly) * r.R + tr.x0 + lx] = tex[h];
		cnt += __popcll(__ballot(w));
	}
	if (__lane_id() == 0)
		atomicAdd(&ntouched, cnt);
	// the tile's count has been read: zero for the next deposit's fill (tiles with no
	// entries return above with it zero already)
	if (lid == 0)
		r.tile_cursor[t] = 0;
	__syncthreads();
	// this tile's diagnostics (summed over tiles on request): touched texels, ranges,
	// serial ranges, sorted, sample / blend / total / window set-up cycles in units of 1024
	if (lid == 0) {
		diag[0] = ntouched;
		diag[1] = windows;
		diag[2] = serial_windows;
		diag[3] = is_sorted ? 1 : 0;
		diag[4] = (int) (csamp >> 10);
		diag[5] = (int) (cblend >> 10);
		diag[6] = (int) ((clock64() - c0) >> 10);
		diag[7] = (int) (cpre >> 10);
	}
}

// Sums the per-tile diagnostics into stats[2..9] (one workgroup):
// [2] ranges, [3] serial ranges, [4] sorted tiles, [5] sample cycles (window set-up and
// samples), [6] blend cycles, [7] total cycles, [8] touched texels, [9] window set-up
// cycles (particle loads, ranks, prefix).
__global__ __launch_bounds__(SCAN_BS) void k_render_diag_sum(int ntiles, const int *tile_diag, long long *stats)
{
	__shared__ long long part[8][SCAN_BS / 64];
	long long acc[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
	for (int t = threadIdx.x; t < ntiles; t += SCAN_BS) {
		const int *d = tile_diag + 8 * (size_t) t;
		acc[0] += d[1];
		acc[1] += d[2];
		acc[2] += d[3];
		acc[3] += (long long) d[4] << 10;
		acc[4] += (long long) d[5] << 10;
		acc[5] += (long long) d[6] << 10;
		acc[6] += d[0];
		acc[7] += (long long) d[7] << 10;
	}
	int w = threadIdx.x / 64;
	for (int k = 0; k < 8; k++) {
		long long v = acc[k];
		for (int off = 32; off > 0; off >>= 1)
			v += __shfl_down(v, off);
		if (__lane_id() == 0)
			part[k][w] = v;
	}
	__syncthreads();
	if (threadIdx.x < 8) {
		long long v = 0;
		for (int i = 0; i < SCAN_BS / 64; i++)
			v += part[threadIdx.x][i];
		stats[2 + threadIdx.x] = v;
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

// reference render.c:253-258
__device__ __forceinline__ unsigned char to8(float v)
{
	if (v <= 0)
		return 0;
	if (v >= 1)
		return 255;
	return (unsigned char) (v * 255.0f + 0.5f);
}

// render_pixel8 (render.c:260-279) of the scaled texel c, in the reference's float
// operations: 4 bytes (un-premultiplied RGBA) without a background, 3 composited over it.
struct Bg {
	int on;
	float r, g, b;
};

__device__ __forceinline__ void pixel8(float c0, float c1, float c2, float c3, const Bg &bg, unsigned char o[4])
{
	float a = c3;
	if (a > 1)
		a = 1;
	if (a < 0)
		a = 0;
	if (!bg.on) {
		if (a > 1e-6f) {
			o[0] = to8(c0 / a);
			o[1] = to8(c1 / a);
			o[2] = to8(c2 / a);
		} else {
			o[0] = o[1] = o[2] = 0;
		}
		o[3] = to8(a);
	} else {
		o[0] = to8(c0 + bg.r * (1 - a));
		o[1] = to8(c1 + bg.g * (1 - a));
		o[2] = to8(c2 + bg.b * (1 - a));
	}
}

static Bg make_bg(const float *bg)
{
	Bg b;
	b.on = bg ? 1 : 0;
	b.r = bg ? bg[0] : 0.0f;
	b.g = bg ? bg[1] : 0.0f;
	b.b = bg ? bg[2] : 0.0f;
	return b;
}

// One texel per thread: c = texel * g (render.c:220-226), then render_pixel8.
__global__ void k_render_faces8(const float4 *buf, size_t ntex, float g, Bg bg, unsigned char *out)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ntex)
		return;
	float4 d = buf[i];
	unsigned char o[4];
	pixel8(d.x * g, d.y * g, d.z * g, d.w * g, bg, o);
	if (!bg.on) {
		((uchar4 *) out)[i] = make_uchar4(o[0], o[1], o[2], o[3]);
	} else {
		out[3 * i] = o[0];
		out[3 * i + 1] = o[1];
		out[3 * i + 2] = o[2];
	}
}

extern "C" void ggbk_render_faces8(const float *buf, size_t ntexels, float g, const float *bg, unsigned char *out)
{
	unsigned nb = (unsigned) ((ntexels + 255) / 256);
	GGBK_LAUNCH(k_render_faces8, nb, 256, 0, (const float4 *) buf, ntexels, g, make_bg(bg), out);
}

// The 4R x 3R cross net (the reference driver's main.c:280-295): one thread per net pixel;
// face f occupies cell (cx[f], cy[f]) = (0,1) (1,1) (2,1) (3,1) (0,0) (0,2); a face pixel
// starts as {0, 0, 0, 255} before render_pixel8 (so it keeps alpha 255 over a background),
// and pixels outside the six cells are {0, 0, 0, 0}.
__global__ void k_render_net8(const float4 *buf, int R, float g, Bg bg, uchar4 *out)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	size_t W = 4 * (size_t) R;
	if (i >= W * 3 * (size_t) R)
		return;
	int X = (int) (i % W), Y = (int) (i / W);
	int cx = X / R, cy = Y / R, x = X - cx * R, y = Y - cy * R;
	int f = cy == 1 ? cx : (cx == 0 ? (cy == 0 ? 4 : 5) : -1);
	unsigned char o[4] = { 0, 0, 0, 0 };
	if (f >= 0) {
		float4 d = buf[((size_t) f * R + y) * R + x];
		o[3] = 255;
		pixel8(d.x * g, d.y * g, d.z * g, d.w * g, bg, o);
	}
	out[i] = make_uchar4(o[0], o[1], o[2], o[3]);
}

extern "C" void ggbk_render_net8(const float *buf, int R, float g, const float *bg, unsigned char *out)
{
	size_t npx = 12 * (size_t) R * R;
	unsigned nb = (unsigned) ((npx + 255) / 256);
	GGBK_LAUNCH(k_render_net8, nb, 256, 0, (const float4 *) buf, R, g, make_bg(bg), (uchar4 *) out);
}

// The equirectangular preview (main.c:297-312): pixel (x, y) looks in the direction of
// longitude (x + 0.5) / W * 2 pi - pi and latitude pi / 2 - (y + 0.5) / H * pi (vec3.h:55-58),
// then render_sample_dir (render.c:228-251): gnomonic texel position on the direction's
// face, clamped, and the float bilinear blend of the four scaled texels, then render_pixel8.
__global__ void k_render_preview8(const float4 *buf, int R, float g, int W, int H, Bg bg, unsigned char *out)
{
	size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (size_t) W * H)
		return;
	int x = (int) (i % (size_t) W), y = (int) (i / (size_t) W);
	double lat = GGB_PI / 2 - (y + 0.5) / H * GGB_PI;
	double lon = (x + 0.5) / W * 2 * GGB_PI - GGB_PI;
	double sla, cla, slo, clo;
	sincos(lat, &sla, &cla);
	sincos(lon, &slo, &clo);
	dv3 d = d3(cla * slo, sla, cla * clo);
	int f = face_of(d);
	double dc, u, v;
	gnomonic(f, d, dc, u, v);
	double px = (u + 1.0) * 0.5 * R - 0.5, py = (1.0 - v) * 0.5 * R - 0.5;
	if (px < 0)
		px = 0;
	if (py < 0)
		py = 0;
	if (px > R - 1)
		px = R - 1;
	if (py > R - 1)
		py = R - 1;
	int x0 = (int) px, y0 = (int) py;
	int x1 = x0 + 1 < R ? x0 + 1 : x0, y1 = y0 + 1 < R ? y0 + 1 : y0;
	float fx = (float) (px - x0), fy = (float) (py - y0);
	const float4 *face = buf + (size_t) f * R * R;
	float4 ta = face[(size_t) y0 * R + x0], tb = face[(size_t) y0 * R + x1];
	float4 tc = face[(size_t) y1 * R + x0], te = face[(size_t) y1 * R + x1];
	float a[4] = { ta.x * g, ta.y * g, ta.z * g, ta.w * g }, b[4] = { tb.x * g, tb.y * g, tb.z * g, tb.w * g };
	float c[4] = { tc.x * g, tc.y * g, tc.z * g, tc.w * g }, e[4] = { te.x * g, te.y * g, te.z * g, te.w * g };
	float o4[4];
#pragma unroll
	for (int k = 0; k < 4; k++)
		o4[k] = (1 - fy) * ((1 - fx) * a[k] + fx * b[k]) + fy * ((1 - fx) * c[k] + fx * e[k]);
	unsigned char o[4];
	pixel8(o4[0], o4[1], o4[2], o4[3], bg, o);
	if (!bg.on) {
		((uchar4 *) out)[i] = make_uchar4(o[0], o[1], o[2], o[3]);
	} else {
		out[3 * i] = o[0];
		out[3 * i + 1] = o[1];
		out[3 * i + 2] = o[2];
	}
}

extern "C" void ggbk_render_preview8(const float *buf, int R, float g, int W, int H, const float *bg,
				     unsigned char *out)
{
	size_t npx = (size_t) W * H;
	unsigned nb = (unsigned) ((npx + 255) / 256);
	GGBK_LAUNCH(k_render_preview8, nb, 256, 0, (const float4 *) buf, R, g, W, H, make_bg(bg), out);
}

extern "C" int ggbk_render_tile(void) { return TILE; }

extern "C" int ggbk_render_bucket(void) { return BUCKET; }

extern "C" int ggbk_render_host_init(ggbk_render *r)
{
	hipEvent_t e1, e2;
	hipStream_t st;
	r->hstats = (long long *) ggbk_host_alloc(2 * sizeof(long long));
	if (!r->hstats)
		return -1;
	if (hipEventCreateWithFlags(&e1, hipEventDisableTiming) != hipSuccess)
		return -1;
	r->stats_ev = (void *) e1;
	if (hipEventCreateWithFlags(&e2, hipEventDisableTiming) != hipSuccess)
		return -1;
	r->fill_ev = (void *) e2;
	if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
		return -1;
	r->copy_stream = (void *) st;
	return 0;
}

extern "C" void ggbk_render_host_free(ggbk_render *r)
{
	if (r->copy_stream)
		(void) hipStreamDestroy((hipStream_t) r->copy_stream);
	if (r->stats_ev)
		(void) hipEventDestroy((hipEvent_t) r->stats_ev);
	if (r->fill_ev)
		(void) hipEventDestroy((hipEvent_t) r->fill_ev);
	ggbk_host_free(r->hstats);
	r->copy_stream = r->stats_ev = r->fill_ev = nullptr;
	r->hstats = nullptr;
}

extern "C" int ggbk_render_scan_blocks(int ntiles) { return (ntiles + SCAN_BS - 1) / SCAN_BS; }

extern "C" int ggbk_render_stats(const ggbk_render *r, long long *out16)
{
	GGBK_LAUNCH(k_render_diag_sum, 1, SCAN_BS, 0, r->ntiles, (const int *) r->tile_diag, r->total);
	return ggbk_d2h(out16, r->total, 16 * sizeof(long long));
}

// Host side of one deposit.  Returns the number of (particle, tile) entries, or -1.
// With reorder, the particles are counted per home tile first and the fill also copies
// them to `out` in that order (the caller adopts it once this returns); without, they
// stay where they are (the storage order does not change the result, only locality).
// The fill writes the tile lists into fixed-size buckets and counts every tile exactly.
// The entry total and the longest list are copied to pinned host memory (on a second
// stream) and the draw is launched at once, in bucket mode; the host then waits for the
// copy only, while the draw runs.  If a tile overflowed its bucket, that draw returns at its start on the
// device, and the counts are scanned into starts, the fill runs again into one packed
// list (grown when the total exceeds its capacity) and the draw runs on that list.
extern "C" long long ggbk_render_deposit(ggbk_render *r, ggbk_parts in, ggbk_parts out, float invg, int reorder)
{
	if (in.n <= 0)
		return 0;
	if (r->tile != TILE || out.n != in.n)
		return -1;
	SplatGeo sg;
	sg.R = r->R;
	sg.ntx = r->ntx;
	sg.rad = r->radius;
	sg.spacing = 0.7 / r->R;
	unsigned pb = ggbk_blocks(in.n, PBS);
	unsigned sb = ggbk_blocks(r->ntiles, SCAN_BS);
	if (in.n > r->nsamp_cap) {
		ggbk_free(r->nsamp);
		r->nsamp = (unsigned short *) ggbk_malloc(sizeof(unsigned short) * (size_t) in.n);
		if (!r->nsamp)
			return -1;
		r->nsamp_cap = in.n;
	}
	int *hcursor = reorder ? r->home_cursor : nullptr;
	if (reorder) {
		// home counts, and their scan (the tile counts are zero here: the scan clears
		// the tile cursors the fill counts into, and the stats)
		ggbk_probe_begin(GGBK_PROBE_RENDER_COUNT);
		if (ggbk_zero(r->tile_count, sizeof(int) * (size_t) r->ntiles) ||
		    ggbk_zero(r->home_count, sizeof(int) * (size_t) r->ntiles))
			return -1;
		GGBK_LAUNCH(k_render_count, pb, PBS, 0, in, sg, r->home_count);
		GGBK_LAUNCH(k_render_scan_local, sb, SCAN_BS, 0, r->ntiles, (const int *) r->tile_count,
			    r->tile_start, r->tile_cursor, (const int *) r->home_count, r->home_start, r->home_cursor,
			    r->scan_part);
		GGBK_LAUNCH(k_render_scan_fix, sb, SCAN_BS, 0, r->ntiles, r->tile_start, r->home_start,
			    (const long long *) r->scan_part, r->total);
		ggbk_probe_end(GGBK_PROBE_RENDER_COUNT);
	}
	// otherwise the tile counts are zero from the last draw, and the fill clears the stats
	ggbk_probe_begin(GGBK_PROBE_RENDER_FILL);
	GGBK_LAUNCH(k_render_fill, pb, PBS, 0, in, out, sg, (const int *) r->tile_start, r->tile_cursor,
		    (const int *) r->home_start, hcursor, r->list, r->cap, r->bucket, r->nsamp, r->total);
	GGBK_LAUNCH(k_render_bucket_stats, sb, SCAN_BS, 0, r->ntiles, (const int *) r->tile_cursor, r->total);
	ggbk_probe_end(GGBK_PROBE_RENDER_FILL);
	// the copy runs on its own stream, after the binning, beside the draw
	hipStream_t cs = (hipStream_t) r->copy_stream;
	if (hipEventRecord((hipEvent_t) r->fill_ev, ggbk_stream()) != hipSuccess ||
	    hipStreamWaitEvent(cs, (hipEvent_t) r->fill_ev, 0) != hipSuccess ||
	    hipMemcpyAsync(r->hstats, r->total, 2 * sizeof(long long), hipMemcpyDeviceToHost, cs) != hipSuccess ||
	    hipEventRecord((hipEvent_t) r->stats_ev, cs) != hipSuccess)
		return -1;
	r->bucket_mode = 1;
	ggbk_probe_begin(GGBK_PROBE_RENDER_DRAW);
	GGBK_LAUNCH(k_render_draw, (unsigned) r->ntiles, DBS, 0, *r, in, sg, invg);
	ggbk_probe_end(GGBK_PROBE_RENDER_DRAW);
	if (hipEventSynchronize((hipEvent_t) r->stats_ev) != hipSuccess)
		return -1;
	long long stats[2] = { r->hstats[0], r->hstats[1] };
	if (stats[0] > INT_MAX / 2)
		return -1;
	if (stats[1] > BUCKET) {
		r->bucket_mode = 0;
		// a tile holds more than a bucket: pack the exact counts into one list
		if (stats[0] > r->cap) {
			long long ncap = stats[0] + stats[0] / 4 + 1024;
			ggbk_free(r->list);
			ggbk_free(r->sort_scratch);
			r->list = (int *) ggbk_malloc(sizeof(int) * (size_t) ncap);
			r->sort_scratch =
				(unsigned long long *) ggbk_malloc(sizeof(unsigned long long) * 2 * (size_t) ncap);
			if (!r->list || !r->sort_scratch)
				return -1;
			r->cap = ncap;
			r->scratch_cap = 2 * ncap;
		}
		if (ggbk_d2d(r->tile_count, r->tile_cursor, sizeof(int) * (size_t) r->ntiles))
			return -1;
		GGBK_LAUNCH(k_render_scan_local, sb, SCAN_BS, 0, r->ntiles, (const int *) r->tile_count,
			    r->tile_start, r->tile_cursor, (const int *) r->home_count, r->home_start, r->home_cursor,
			    r->scan_part);
		GGBK_LAUNCH(k_render_scan_fix, sb, SCAN_BS, 0, r->ntiles, r->tile_start, r->home_start,
			    (const long long *) r->scan_part, r->total);
		GGBK_LAUNCH(k_render_fill, pb, PBS, 0, in, out, sg, (const int *) r->tile_start, r->tile_cursor,
			    (const int *) r->home_start, hcursor, r->list, r->cap, (int *) nullptr, r->nsamp,
			    r->total);
		ggbk_probe_begin(GGBK_PROBE_RENDER_DRAW);
		GGBK_LAUNCH(k_render_draw, (unsigned) r->ntiles, DBS, 0, *r, in, sg, invg);
		ggbk_probe_end(GGBK_PROBE_RENDER_DRAW);
	}
	if (ggbk_check("render_deposit"))
		return -1;
	return stats[0];
}
