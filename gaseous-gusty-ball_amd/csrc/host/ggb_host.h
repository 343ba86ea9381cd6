/* Internal declarations shared by the host C modules. */
#ifndef GGB_HOST_H
#define GGB_HOST_H

#include <stddef.h>

#include "cubesphere.h"
#include "fluid.h"
#include "particles.h"
#include "../kern/ggbk.h"

/* Device copy of a grid's tables (column-major ELL), uploaded on first use and
 * released by cs_free. */
struct ggb_grid_dev {
	ggbk_grid g;
	void *bufs[13];
	int nbufs;
};

/* Returns the device view of cs, uploading it on first use; aborts on failure. */
ggbk_grid ggb_grid_dev(const CubeSphere *cs);
void ggb_grid_dev_release(CubeSphere *cs);

/* Device allocation that aborts with a message on failure (the reference's in-step
 * allocation failure behaviour, fluid.c:8-16). */
void *ggb_dmalloc(size_t bytes, const char *what);
void ggb_die(const char *fmt, ...);
int ggb_ensure_device(void);
/* Whether fluid_step and particles_advect refresh the host mirrors (ggb_set_sync_mirrors,
 * or GGB_SYNC_MIRRORS=1 in the environment). */
int ggb_sync_mirrors(void);

/* Live fluids, so particles_advect can map F.vel to the device velocity. */
const Fluid *ggb_fluid_of_vel(const float *vel);
/* device velocity and its element type of a live fluid */
const void *ggb_fluid_dev_vel(const Fluid *F, int *T);

/* A particle set's current device arrays, and those the renderer writes the next
 * (tile) storage order into; ggb_particles_dev_swap adopts the written order. */
void ggb_particles_dev_parts(const Particles *P, ggbk_parts *cur, ggbk_parts *nxt);
void ggb_particles_dev_swap(const Particles *P);

#endif
