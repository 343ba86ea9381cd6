/*
 * Particles, host side (reference src/particles.c): device arrays in double, float
 * host mirrors, and the velocity-handle lookup that lets particles_advect read a live
 * Fluid's device velocity directly.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ggb_ext.h"
#include "ggb_host.h"
#include "particles.h"

struct ggb_particles_dev {
	double *pos, *prev;
	float *col;
	int round32;
	void *vel_tmp;     /* staging for host velocity arrays */
	size_t vel_tmp_bytes;
};

static void mirror_positions(Particles *P)
{
	size_t n = (size_t) P->n;
	double *tmp = malloc(sizeof(double) * 3 * n);
	if (!tmp)
		ggb_die("particles: out of memory");
	if (ggbk_d2h(tmp, P->dev->pos, 24 * n))
		ggb_die("particles: %s", ggbk_last_error());
	for (size_t i = 0; i < 3 * n; i++)
		P->pos[i] = (float) tmp[i];
	if (ggbk_d2h(tmp, P->dev->prev, 24 * n))
		ggb_die("particles: %s", ggbk_last_error());
	for (size_t i = 0; i < 3 * n; i++)
		P->prev[i] = (float) tmp[i];
	free(tmp);
}

int particles_init(Particles *P, int n)
{
	memset(P, 0, sizeof(*P));
	if (n < 0)
		return -1;
	if (ggb_ensure_device()) {
		fprintf(stderr, "particles: no GPU available\n");
		return -1;
	}
	P->n = n;
	P->pos = malloc(sizeof(float) * 3 * (size_t) (n ? n : 1));
	P->prev = malloc(sizeof(float) * 3 * (size_t) (n ? n : 1));
	P->col = malloc(sizeof(float) * 4 * (size_t) (n ? n : 1));
	struct ggb_particles_dev *D = calloc(1, sizeof(*D));
	if (!P->pos || !P->prev || !P->col || !D) {
		fprintf(stderr, "particles: out of memory\n");
		return -1;
	}
	P->dev = D;
	D->pos = ggbk_malloc(24 * (size_t) n);
	D->prev = ggbk_malloc(24 * (size_t) n);
	D->col = ggbk_malloc(16 * (size_t) n);
	if (!D->pos || !D->prev || !D->col) {
		fprintf(stderr, "particles: out of device memory\n");
		return -1;
	}
	if (n > 0) {
		ggbk_particles_init(n, D->pos, D->prev, D->col, 0);
		mirror_positions(P);
		if (ggbk_d2h(P->col, D->col, 16 * (size_t) n))
			return -1;
	}
	return 0;
}

void particles_free(Particles *P)
{
	if (P->dev) {
		ggb_sync();
		ggbk_free(P->dev->pos);
		ggbk_free(P->dev->prev);
		ggbk_free(P->dev->col);
		ggbk_free(P->dev->vel_tmp);
		free(P->dev);
	}
	free(P->pos);
	free(P->prev);
	free(P->col);
	memset(P, 0, sizeof(*P));
}

void particles_color_from_image(Particles *P, const float *img, int w, int h)
{
	if (P->n <= 0 || w <= 0 || h <= 0)
		return;
	size_t bytes = sizeof(float) * 4 * (size_t) w * h;
	float *dimg = ggb_dmalloc(bytes, "particles image");
	if (ggbk_h2d(dimg, img, bytes))
		ggb_die("particles: %s", ggbk_last_error());
	ggbk_particles_color(P->n, P->dev->pos, P->dev->col, dimg, w, h);
	if (ggbk_d2h(P->col, P->dev->col, 16 * (size_t) P->n))
		ggb_die("particles: %s", ggbk_last_error());
	ggbk_free(dimg);
}

void particles_advect(Particles *P, const CubeSphere *cs, const float *vel, double dt)
{
	if (P->n <= 0)
		return;
	struct ggb_particles_dev *D = P->dev;
	ggbk_grid g = ggb_grid_dev(cs);
	const Fluid *F = ggb_fluid_of_vel(vel);
	const void *dvel;
	int T;
	if (F && F->cs->ncells == cs->ncells) {
		dvel = ggb_fluid_dev_vel(F, &T);
	} else {
		size_t bytes = sizeof(float) * 3 * (size_t) cs->ncells;
		if (D->vel_tmp_bytes < bytes) {
			ggb_sync();
			ggbk_free(D->vel_tmp);
			D->vel_tmp = ggb_dmalloc(bytes, "particles velocity");
			D->vel_tmp_bytes = bytes;
		}
		if (ggbk_h2d(D->vel_tmp, vel, bytes))
			ggb_die("particles: %s", ggbk_last_error());
		dvel = D->vel_tmp;
		T = GGBK_F32;
	}
	ggbk_particles_advect(g, P->n, D->pos, D->prev, dvel, T, dt, D->round32);
}

void particles_sync_host(Particles *P)
{
	if (P->n > 0)
		mirror_positions(P);
}

int particles_get_f64(const Particles *P, double *pos, double *prev)
{
	size_t n = (size_t) P->n;
	if (pos && ggbk_d2h(pos, P->dev->pos, 24 * n))
		return -1;
	if (prev && ggbk_d2h(prev, P->dev->prev, 24 * n))
		return -1;
	return 0;
}

int particles_set_f64(Particles *P, const double *pos, const double *prev)
{
	size_t n = (size_t) P->n;
	if (pos && ggbk_h2d(P->dev->pos, pos, 24 * n))
		return -1;
	if (prev && ggbk_h2d(P->dev->prev, prev, 24 * n))
		return -1;
	mirror_positions(P);
	return 0;
}

void particles_set_round32(Particles *P, int on) { P->dev->round32 = on ? 1 : 0; }

/* device pointers for the renderer */
void ggb_particles_dev_arrays(const Particles *P, const double **prev, const double **pos, const float **col)
{
	*prev = P->dev->prev;
	*pos = P->dev->pos;
	*col = P->dev->col;
}
