/*
 * Particles, host side (reference src/particles.c): device arrays in double, float
 * host mirrors, and the velocity-handle lookup that lets particles_advect read a live
 * Fluid's device velocity directly.
 *
 * The device arrays are in storage order: slot s holds particle pid[s].  Each render
 * deposit rewrites them in tile order into the second set of arrays, which then becomes
 * current; every host view (mirrors, get/set) is in particle index order, as in the
 * reference.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ggb_ext.h"
#include "ggb_host.h"
#include "particles.h"

struct ggb_particles_dev {
	ggbk_parts cur, nxt;
	int round32;
	void *vel_tmp;     /* staging for host velocity arrays */
	size_t vel_tmp_bytes;
};

static int parts_alloc(ggbk_parts *p, int n)
{
	memset(p, 0, sizeof(*p));
	p->n = n;
	p->pos = ggbk_malloc(24 * (size_t) n);
	p->prev = ggbk_malloc(24 * (size_t) n);
	p->col = ggbk_malloc(16 * (size_t) n);
	p->pid = ggbk_malloc(4 * (size_t) n);
	return p->pos && p->prev && p->col && p->pid ? 0 : -1;
}

static void parts_free(ggbk_parts *p)
{
	ggbk_free(p->pos);
	ggbk_free(p->prev);
	ggbk_free(p->col);
	ggbk_free(p->pid);
	memset(p, 0, sizeof(*p));
}

static int *download_pid(const Particles *P)
{
	int *pid = malloc(sizeof(int) * (size_t) P->n);
	if (!pid)
		ggb_die("particles: out of memory");
	if (ggbk_d2h(pid, P->dev->cur.pid, 4 * (size_t) P->n))
		ggb_die("particles: %s", ggbk_last_error());
	return pid;
}

/* device vec3 array in storage order -> host array in index order */
static int gather_d(const Particles *P, const int *pid, const double *dsrc, double *out)
{
	size_t n = (size_t) P->n;
	double *tmp = malloc(sizeof(double) * 3 * n);
	if (!tmp)
		return -1;
	if (ggbk_d2h(tmp, dsrc, 24 * n)) {
		free(tmp);
		return -1;
	}
	for (size_t s = 0; s < n; s++)
		memcpy(out + 3 * (size_t) pid[s], tmp + 3 * s, 24);
	free(tmp);
	return 0;
}

static void mirror_positions(Particles *P)
{
	size_t n = (size_t) P->n;
	int *pid = download_pid(P);
	double *tmp = malloc(sizeof(double) * 3 * n);
	if (!tmp)
		ggb_die("particles: out of memory");
	if (gather_d(P, pid, P->dev->cur.pos, tmp))
		ggb_die("particles: %s", ggbk_last_error());
	for (size_t i = 0; i < 3 * n; i++)
		P->pos[i] = (float) tmp[i];
	if (gather_d(P, pid, P->dev->cur.prev, tmp))
		ggb_die("particles: %s", ggbk_last_error());
	for (size_t i = 0; i < 3 * n; i++)
		P->prev[i] = (float) tmp[i];
	free(tmp);
	free(pid);
}

static void mirror_colors(Particles *P)
{
	size_t n = (size_t) P->n;
	int *pid = download_pid(P);
	float *tmp = malloc(sizeof(float) * 4 * n);
	if (!tmp)
		ggb_die("particles: out of memory");
	if (ggbk_d2h(tmp, P->dev->cur.col, 16 * n))
		ggb_die("particles: %s", ggbk_last_error());
	for (size_t s = 0; s < n; s++)
		memcpy(P->col + 4 * (size_t) pid[s], tmp + 4 * s, 16);
	free(tmp);
	free(pid);
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
	if (parts_alloc(&D->cur, n) || parts_alloc(&D->nxt, n)) {
		fprintf(stderr, "particles: out of device memory\n");
		return -1;
	}
	if (n > 0) {
		ggbk_particles_init(D->cur, 0);
		mirror_positions(P);
		mirror_colors(P);
	}
	return 0;
}

void particles_free(Particles *P)
{
	if (P->dev) {
		ggb_sync();
		parts_free(&P->dev->cur);
		parts_free(&P->dev->nxt);
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
	ggbk_particles_color(P->n, P->dev->cur.pos, P->dev->cur.col, dimg, w, h);
	mirror_colors(P);
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
	ggbk_particles_advect(g, P->n, D->cur.pos, D->cur.prev, dvel, T, dt, D->round32);
	if (ggb_sync_mirrors())
		particles_sync_host(P);
}

void particles_sync_host(Particles *P)
{
	if (P->n > 0)
		mirror_positions(P);
}

int particles_get_f64(const Particles *P, double *pos, double *prev)
{
	if (P->n <= 0)
		return 0;
	int *pid = download_pid(P);
	int rc = 0;
	if (pos && gather_d(P, pid, P->dev->cur.pos, pos))
		rc = -1;
	if (!rc && prev && gather_d(P, pid, P->dev->cur.prev, prev))
		rc = -1;
	free(pid);
	return rc;
}

/* host array in index order -> device array in storage order */
static int scatter_d(const Particles *P, const int *pid, const double *src, double *ddst)
{
	size_t n = (size_t) P->n;
	double *tmp = malloc(sizeof(double) * 3 * n);
	if (!tmp)
		return -1;
	for (size_t s = 0; s < n; s++)
		memcpy(tmp + 3 * s, src + 3 * (size_t) pid[s], 24);
	int rc = ggbk_h2d(ddst, tmp, 24 * n);
	free(tmp);
	return rc;
}

int particles_set_f64(Particles *P, const double *pos, const double *prev)
{
	if (P->n <= 0)
		return 0;
	int *pid = download_pid(P);
	int rc = 0;
	if (pos && scatter_d(P, pid, pos, P->dev->cur.pos))
		rc = -1;
	if (!rc && prev && scatter_d(P, pid, prev, P->dev->cur.prev))
		rc = -1;
	free(pid);
	if (!rc)
		mirror_positions(P);
	return rc;
}

void particles_set_round32(Particles *P, int on) { P->dev->round32 = on ? 1 : 0; }

/* the renderer's view: current arrays and the arrays of the next storage order */
void ggb_particles_dev_parts(const Particles *P, ggbk_parts *cur, ggbk_parts *nxt)
{
	*cur = P->dev->cur;
	*nxt = P->dev->nxt;
}

/* adopts the storage order the renderer wrote (the caller's struct is unchanged; only
 * the device-side bookkeeping moves) */
void ggb_particles_dev_swap(const Particles *P)
{
	ggbk_parts t = P->dev->cur;
	P->dev->cur = P->dev->nxt;
	P->dev->nxt = t;
}
