/* Passive coloured tracer particles on the unit sphere, advected on one GPU.
 *
 * Drop-in for the reference's src/particles.h:7-21.  Positions live on the device in
 * double; pos / prev / col here are HOST MIRRORS, valid after particles_init,
 * particles_color_from_image and particles_sync_host() (ggb_ext.h).
 */
#ifndef GGB_PARTICLES_H
#define GGB_PARTICLES_H

#include "cubesphere.h"

#ifdef __cplusplus
extern "C" {
#endif

struct ggb_particles_dev;

typedef struct {
	int n;
	float *pos;  /* [n*3] host mirror: current positions */
	float *prev; /* [n*3] host mirror: positions before the last advection */
	float *col;  /* [n*4] RGBA colours (fixed after particles_color_from_image) */
	struct ggb_particles_dev *dev;
} Particles;

/* reference particles.h:14-21 / particles.c:8-74 */
int particles_init(Particles *P, int n);
void particles_free(Particles *P);
void particles_color_from_image(Particles *P, const float *img, int w, int h);
/* RK2 midpoint advection along great circles.  `vel` may be a live Fluid's F.vel (the
 * device velocity is used directly) or any host array of ncells*3 floats. */
void particles_advect(Particles *P, const CubeSphere *cs, const float *vel, double dt);

#ifdef __cplusplus
}
#endif

#endif
