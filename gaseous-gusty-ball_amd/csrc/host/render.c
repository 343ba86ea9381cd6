/*
 * Trail renderer, host side (reference src/render.c): the global decay factor g and
 * its renormalisation, the device buffer, and the host readers.  The readers
 * (render_texel, render_sample_dir) refresh the host mirror once per change of the
 * device buffer; the refresh is guarded so concurrent OpenMP readers are safe.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ggb_ext.h"
#include "ggb_host.h"
#include "render.h"

struct ggb_render_dev {
	ggbk_render k;
	long long last_entries;
	unsigned long gen_dev;  /* bumped by every change of the device buffer */
	unsigned long gen_host; /* generation the mirror holds */
	pthread_mutex_t mu;
	unsigned char *faces8;  /* device 8-bit faces (render_faces8) */
	size_t faces8_bytes;
	unsigned long ndeposits;
};
#ifndef GGB_REORDER_EVERY
#define GGB_REORDER_EVERY 16
#endif

int render_init(Render *r, int R, float radius, float alpha)
{
	memset(r, 0, sizeof(*r));
	if (R < 1)
		return -1;
	if (ggb_ensure_device()) {
		fprintf(stderr, "render: no GPU available\n");
		return -1;
	}
	r->R = R;
	r->radius = radius < 0.75f ? 0.75f : radius;
	r->alpha = alpha;
	r->g = 1.0;
	size_t count = (size_t) 6 * R * R * 4;
	r->buf = calloc(count, sizeof(float));
	struct ggb_render_dev *D = calloc(1, sizeof(*D));
	if (!r->buf || !D) {
		fprintf(stderr, "render: out of memory\n");
		return -1;
	}
	r->dev = D;
	pthread_mutex_init(&D->mu, NULL);
	/* the reference's strip binning parameters (render.c:23-30), kept for callers */
	int sh = 32;
	if (sh < (int) ceilf(r->radius) + 1)
		sh = (int) ceilf(r->radius) + 1;
	if (sh > R)
		sh = R;
	r->strip_h = sh;
	r->nstrips = (R + sh - 1) / sh;
	r->nbins = 6 * r->nstrips;

	ggbk_render *k = &D->k;
	k->R = R;
	k->tile = ggbk_render_tile();
	k->ntx = (R + k->tile - 1) / k->tile;
	k->ntiles = 6 * k->ntx * k->ntx;
	k->radius = r->radius;
	k->alpha = alpha;
	k->buf = ggbk_malloc(sizeof(float) * count);
	k->tile_count = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->tile_start = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->tile_cursor = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->home_count = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->home_start = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->home_cursor = ggbk_malloc(sizeof(int) * ((size_t) k->ntiles + 1));
	k->total = ggbk_malloc(sizeof(long long) * 16);
	k->tile_diag = ggbk_malloc(sizeof(int) * 8 * (size_t) k->ntiles);
	k->scan_part = ggbk_malloc(sizeof(long long) * 3 * (size_t) ggbk_render_scan_blocks(k->ntiles));
	k->bucket = ggbk_malloc(sizeof(int) * (size_t) ggbk_render_bucket() * (size_t) k->ntiles);
	k->bucket_mode = 0;
	k->cap = 0;
	k->list = NULL;
	k->nsamp = NULL;
	k->nsamp_cap = 0;
	k->sort_scratch = NULL;
	k->hstats = NULL;
	k->stats_ev = NULL;
	k->fill_ev = NULL;
	k->copy_stream = NULL;
	if (!k->buf || !k->tile_count || !k->tile_start || !k->tile_cursor || !k->home_count || !k->home_start ||
	    !k->home_cursor || !k->total || !k->tile_diag ||
	    !k->scan_part || !k->bucket) {
		fprintf(stderr, "render: out of device memory\n");
		return -1;
	}
	/* the draw leaves the tile counts at zero for the next deposit's fill: start there */
	if (ggbk_zero(k->tile_cursor, sizeof(int) * ((size_t) k->ntiles + 1))) {
		fprintf(stderr, "render: %s\n", ggbk_last_error());
		return -1;
	}
	if (ggbk_render_host_init(k)) {
		fprintf(stderr, "render: out of pinned host memory\n");
		return -1;
	}
	if (ggbk_zero(k->buf, sizeof(float) * count) || ggbk_zero(k->tile_diag, sizeof(int) * 8 * (size_t) k->ntiles) ||
	    ggbk_zero(k->total, sizeof(long long) * 16))
		return -1;
	D->gen_dev = 1;
	D->gen_host = 1; /* both zero */
	return 0;
}

void render_free(Render *r)
{
	struct ggb_render_dev *D = r->dev;
	if (D) {
		ggb_sync();
		ggbk_free(D->k.buf);
		ggbk_free(D->k.tile_count);
		ggbk_free(D->k.tile_start);
		ggbk_free(D->k.tile_cursor);
		ggbk_free(D->k.home_count);
		ggbk_free(D->k.home_start);
		ggbk_free(D->k.home_cursor);
		ggbk_free(D->k.total);
		ggbk_free(D->k.tile_diag);
		ggbk_free(D->k.scan_part);
		ggbk_free(D->k.bucket);
		ggbk_free(D->faces8);
		ggbk_free(D->k.list);
		ggbk_free(D->k.nsamp);
		ggbk_free(D->k.sort_scratch);
		ggbk_render_host_free(&D->k);
		pthread_mutex_destroy(&D->mu);
		free(D);
	}
	free(r->buf);
	memset(r, 0, sizeof(*r));
}

/* reference render.c:55-66 */
void render_fade(Render *r, double fade)
{
	r->g *= (1.0 - fade);
	if (r->g < 1e-4) {
		size_t n = (size_t) 6 * r->R * r->R * 4;
		ggbk_render_scale(r->dev->k.buf, n, (float) r->g);
		r->g = 1.0;
		r->dev->gen_dev++;
	}
}

void render_deposit(Render *r, const Particles *P)
{
	ggbk_parts cur, nxt;
	if (P->n <= 0)
		return;
	ggb_particles_dev_parts(P, &cur, &nxt);
	float invg = (float) (1.0 / r->g);
	/* particles move a few texels per step: tile order is restored every few deposits */
	int reorder = r->dev->ndeposits++ % GGB_REORDER_EVERY == 0;
	long long e = ggbk_render_deposit(&r->dev->k, cur, nxt, invg, reorder);
	if (e < 0)
		ggb_die("render: deposit failed (%s)", ggbk_last_error());
	if (reorder)
		ggb_particles_dev_swap(P);
	r->dev->last_entries = e;
	r->dev->gen_dev++;
}

long long render_last_entries(const Render *r) { return r->dev->last_entries; }

int render_last_stats(const Render *r, long long *out16) { return ggbk_render_stats(&r->dev->k, out16); }

static void refresh(const Render *r)
{
	struct ggb_render_dev *D = r->dev;
	if (__atomic_load_n(&D->gen_host, __ATOMIC_ACQUIRE) == D->gen_dev)
		return;
	pthread_mutex_lock(&D->mu);
	if (D->gen_host != D->gen_dev) {
		size_t bytes = sizeof(float) * 6 * (size_t) r->R * r->R * 4;
		if (ggbk_d2h(r->buf, D->k.buf, bytes))
			ggb_die("render: %s", ggbk_last_error());
		__atomic_store_n(&D->gen_host, D->gen_dev, __ATOMIC_RELEASE);
	}
	pthread_mutex_unlock(&D->mu);
}

void render_sync_host(Render *r) { refresh(r); }

/* reference render.c:220-226 */
void render_texel(const Render *r, int face, int x, int y, float out[4])
{
	refresh(r);
	const float *d = r->buf + (((size_t) face * r->R + y) * r->R + x) * 4;
	float g = (float) r->g;
	for (int c = 0; c < 4; c++)
		out[c] = d[c] * g;
}

/* reference render.c:228-251 */
void render_sample_dir(const Render *r, double dx, double dy, double dz, float out[4])
{
	vec3 d = v3(dx, dy, dz);
	int f = cs_face_of(d);
	double dc = v3_dot(d, cs_face_center[f]);
	double u = v3_dot(d, cs_face_right[f]) / dc;
	double v = v3_dot(d, cs_face_up[f]) / dc;
	int R = r->R;
	double px = (u + 1.0) * 0.5 * R - 0.5, py = (1.0 - v) * 0.5 * R - 0.5;
	if (px < 0) px = 0;
	if (py < 0) py = 0;
	if (px > R - 1) px = R - 1;
	if (py > R - 1) py = R - 1;
	int x0 = (int) px, y0 = (int) py;
	int x1 = x0 + 1 < R ? x0 + 1 : x0, y1 = y0 + 1 < R ? y0 + 1 : y0;
	float fx = (float) (px - x0), fy = (float) (py - y0);
	float a[4], b[4], c[4], e[4];
	render_texel(r, f, x0, y0, a);
	render_texel(r, f, x1, y0, b);
	render_texel(r, f, x0, y1, c);
	render_texel(r, f, x1, y1, e);
	for (int k = 0; k < 4; k++)
		out[k] = (1 - fy) * ((1 - fx) * a[k] + fx * b[k]) + fy * ((1 - fx) * c[k] + fx * e[k]);
}

static inline unsigned char to8(float v)
{
	if (v <= 0)
		return 0;
	if (v >= 1)
		return 255;
	return (unsigned char) (v * 255.0f + 0.5f);
}

/* reference render.c:260-279 */
void render_pixel8(const float c[4], const float *bg, unsigned char *out)
{
	float a = c[3];
	if (a > 1)
		a = 1;
	if (a < 0)
		a = 0;
	if (!bg) {
		if (a > 1e-6f) {
			out[0] = to8(c[0] / a);
			out[1] = to8(c[1] / a);
			out[2] = to8(c[2] / a);
		} else {
			out[0] = out[1] = out[2] = 0;
		}
		out[3] = to8(a);
	} else {
		out[0] = to8(c[0] + bg[0] * (1 - a));
		out[1] = to8(c[1] + bg[1] * (1 - a));
		out[2] = to8(c[2] + bg[2] * (1 - a));
	}
}

/* Device staging for the 8-bit outputs (grown on demand). */
static unsigned char *out8_buf(struct ggb_render_dev *D, size_t bytes)
{
	if (D->faces8_bytes < bytes) {
		ggbk_free(D->faces8);
		D->faces8 = ggb_dmalloc(bytes, "render 8-bit output");
		D->faces8_bytes = bytes;
	}
	return D->faces8;
}

/* The whole buffer through render_texel + render_pixel8, converted on the device: the
 * copy back is the 8-bit faces (a quarter or less of the float buffer). */
void render_faces8(Render *r, const float *bg, unsigned char *out)
{
	struct ggb_render_dev *D = r->dev;
	int ch = bg ? 3 : 4;
	size_t ntex = 6 * (size_t) r->R * r->R, bytes = ntex * ch;
	unsigned char *d = out8_buf(D, ntex * 4);
	ggbk_render_faces8(D->k.buf, ntex, (float) r->g, bg, d);
	if (ggbk_d2h(out, d, bytes))
		ggb_die("render: %s", ggbk_last_error());
}

/* main.c:280-295 on the device */
void render_net8(Render *r, const float *bg, unsigned char *out)
{
	struct ggb_render_dev *D = r->dev;
	size_t bytes = 12 * (size_t) r->R * r->R * 4;
	unsigned char *d = out8_buf(D, bytes);
	ggbk_render_net8(D->k.buf, r->R, (float) r->g, bg, d);
	if (ggbk_d2h(out, d, bytes))
		ggb_die("render: %s", ggbk_last_error());
}

/* main.c:297-312 on the device */
int render_preview8(Render *r, const float *bg, int W, int H, unsigned char *out)
{
	struct ggb_render_dev *D = r->dev;
	if (W < 1 || H < 1)
		return -1;
	size_t bytes = (size_t) W * H * (bg ? 3 : 4);
	unsigned char *d = out8_buf(D, (size_t) W * H * 4);
	ggbk_render_preview8(D->k.buf, r->R, (float) r->g, W, H, bg, d);
	if (ggbk_d2h(out, d, bytes))
		ggb_die("render: %s", ggbk_last_error());
	return 0;
}
