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

/* A Galerkin coarse operator of the mixed-precision hierarchy on the host (galerkin.c):
 * (K x)_i = sum_k w[i][k] (x_i - x[nbr[i][k]]), GGBK_WNB slots per cell.  A regular cell
 * has its slots at the fixed 5x5 offsets (ggbk_wide); any other lists its nn neighbours by
 * ascending index, self-padded. */
typedef struct {
	int n, N;
	int *nbr;            /* [n][GGBK_WNB] */
	double *w;           /* [n][GGBK_WNB] the stored half floats' values, zero-padded */
	unsigned short *wh;  /* [n][GGBK_WNB] their bits */
	unsigned char *nn;   /* [n] slots in use (GGBK_WNB for a regular cell) */
	unsigned char *reg;  /* [n] */
	float *dinv;         /* [n] 1 / sum_k w[i][k] */
} GalOp;

/* K_c = R K_f P into *out for the coarse grid of face size Nc: K_f the Voronoi operator
 * of cs (gf == NULL) or the Galerkin operator gf; P given per fine cell by 4 taps and
 * weights, R = P^T as rows (rstart, ridx, rw).  Returns -1 (nothing allocated) when a
 * row does not fit the slots. */
int ggb_galerkin(const CubeSphere *cs, const GalOp *gf, const int *ptap, const float *pw, const int *rstart,
		 const int *ridx, const float *rw, int Nc, GalOp *out);
void ggb_galop_free(GalOp *g);
unsigned short ggb_f16_bits(double v);
double ggb_f16_value(unsigned short h);

#endif
