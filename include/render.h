/* Fading, alpha-blended particle-trail renderer into 6 cubemap faces, on one GPU.
 *
 * Drop-in for the reference's src/render.h:17-54.  The buffer holds premultiplied RGBA
 * divided by a global decay factor g (fading is O(1) per step; the buffer is rescaled
 * when g drops below 1e-4).  Splats are blended, per texel, in (particle index, sample)
 * order -- the reference's order -- by binning them per face tile on the device.
 *
 * `buf` is a HOST MIRROR: render_texel, render_sample_dir and render_sync_host()
 * refresh it lazily (once per change, safe to call from several threads).
 */
#ifndef GGB_RENDER_H
#define GGB_RENDER_H

#include <stddef.h>
#include <stdint.h>

#include "particles.h"

#ifdef __cplusplus
extern "C" {
#endif

struct ggb_render_dev;

typedef struct {
	int R;
	float *buf;      /* [6*R*R*4] host mirror of the scaled premultiplied RGBA */
	double g;        /* global decay factor */
	float radius;    /* splat radius in texels (clamped to >= 0.75) */
	float alpha;     /* per-splat opacity */
	int strip_h, nstrips, nbins; /* the reference's strip binning, for information */
	struct ggb_render_dev *dev;
} Render;

/* reference render.h:44-54 */
int render_init(Render *r, int R, float radius, float alpha);
void render_free(Render *r);
void render_fade(Render *r, double fade);
void render_deposit(Render *r, const Particles *P);
void render_texel(const Render *r, int face, int x, int y, float out[4]);
void render_sample_dir(const Render *r, double dx, double dy, double dz, float out[4]);
void render_pixel8(const float c[4], const float *bg, unsigned char *out);

#ifdef __cplusplus
}
#endif

#endif
