/*
 * Cubed-sphere geometry (host side): cell centres, exact spherical-Voronoi stencil,
 * least-squares gradient coefficients, the seam ghost table, and bilinear sampling.
 *
 * This is setup, not per-step work: it runs once per grid level on the host (OpenMP),
 * and the tables are uploaded to the GPU when a Fluid first uses the grid.  The tables
 * must match the reference's bit for bit (every stencil coefficient and neighbour
 * order feeds the per-step kernels), so the construction follows the reference's
 * arithmetic exactly:
 *   - centres: equiangular face angles, tan, normalise by reciprocal length
 *     (reference cubesphere.c:22-27, 53-56, 254-258);
 *   - Voronoi cell: half-plane clipping of a 8h square in the gnomonic plane at the
 *     centre, candidates from the 5x5 block in (dj, di) row order with seam ghosts
 *     resolved to real cells (cubesphere.c:130-225);
 *   - links sorted by edge length, longest first, ties kept in discovery order (the
 *     stable merge sort glibc's qsort uses for these arrays), cut at 1e-4 h
 *     (cubesphere.c:227-283);
 *   - reciprocity: keep mutual links, average the two edge lengths (cubesphere.c:287-308);
 *   - w = len/dist, gradient coefficients from the 2x2 weighted normal equations
 *     (cubesphere.c:310-351).
 * Reductions (total area, counters) are summed in cell order, so the result does not
 * depend on the thread count.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cubesphere.h"
#include "ggb_host.h"

const vec3 cs_face_center[6] = { { 0, 0, 1 }, { 1, 0, 0 }, { 0, 0, -1 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 } };
const vec3 cs_face_right[6] = { { 1, 0, 0 }, { 0, 0, -1 }, { -1, 0, 0 }, { 0, 0, 1 }, { 1, 0, 0 }, { 1, 0, 0 } };
const vec3 cs_face_up[6] = { { 0, 1, 0 }, { 0, 1, 0 }, { 0, 1, 0 }, { 0, 1, 0 }, { 0, 0, -1 }, { 0, 0, 1 } };

int cs_face_of(vec3 d)
{
	double ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
	if (ax >= ay && ax >= az)
		return d.x > 0 ? 1 : 3;
	if (ay >= az)
		return d.y > 0 ? 4 : 5;
	return d.z > 0 ? 0 : 2;
}

vec3 cs_face_ab_to_dir(int f, double a, double b)
{
	vec3 off = v3_add(v3_scale(cs_face_right[f], tan(a)), v3_scale(cs_face_up[f], tan(b)));
	return v3_normalize(v3_add(cs_face_center[f], off));
}

void cs_locate(const CubeSphere *cs, vec3 d, int *f, double *x, double *y)
{
	int face = cs_face_of(d);
	double dc = v3_dot(d, cs_face_center[face]);
	double u = v3_dot(d, cs_face_right[face]) / dc;
	double v = v3_dot(d, cs_face_up[face]) / dc;
	*f = face;
	*x = (atan(u) + M_PI / 4) / cs->h - 0.5;
	*y = (atan(v) + M_PI / 4) / cs->h - 0.5;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int cs_cell_of(const CubeSphere *cs, vec3 d)
{
	int f;
	double x, y;
	cs_locate(cs, d, &f, &x, &y);
	int i = clampi((int) floor(x + 0.5), 0, cs->N - 1);
	int j = clampi((int) floor(y + 0.5), 0, cs->N - 1);
	return cs_index(cs, f, i, j);
}

/* centre direction of cell (i, j) of face f; i, j may lie outside [0, N) */
static vec3 centre_dir(const CubeSphere *cs, int f, int i, int j)
{
	return cs_face_ab_to_dir(f, -M_PI / 4 + (i + 0.5) * cs->h, -M_PI / 4 + (j + 0.5) * cs->h);
}

/* real cell for (f, i, j): itself inside the face, else the cell holding the extended
 * centre (a seam ghost) */
static int resolve_cell(const CubeSphere *cs, int f, int i, int j)
{
	if (i >= 0 && j >= 0 && i < cs->N && j < cs->N)
		return cs_index(cs, f, i, j);
	return cs_cell_of(cs, centre_dir(cs, f, i, j));
}

static inline int ghost_at(const CubeSphere *cs, int f, int i, int j)
{
	int P = cs->N + 2;
	return cs->ghost[((size_t) f * P + (j + 1)) * P + (i + 1)];
}

void cs_taps(const CubeSphere *cs, vec3 p, int idx[4], double wt[4])
{
	int f, N = cs->N;
	double x, y;
	cs_locate(cs, p, &f, &x, &y);
	int i0 = (int) floor(x), j0 = (int) floor(y);
	double fx = x - i0, fy = y - j0;
	if (i0 >= 0 && j0 >= 0 && i0 + 1 < N && j0 + 1 < N) {
		int b = cs_index(cs, f, i0, j0);
		idx[0] = b;
		idx[1] = b + 1;
		idx[2] = b + N;
		idx[3] = b + N + 1;
	} else if (cs->ghost) {
		idx[0] = ghost_at(cs, f, i0, j0);
		idx[1] = ghost_at(cs, f, i0 + 1, j0);
		idx[2] = ghost_at(cs, f, i0, j0 + 1);
		idx[3] = ghost_at(cs, f, i0 + 1, j0 + 1);
	} else {
		idx[0] = resolve_cell(cs, f, i0, j0);
		idx[1] = resolve_cell(cs, f, i0 + 1, j0);
		idx[2] = resolve_cell(cs, f, i0, j0 + 1);
		idx[3] = resolve_cell(cs, f, i0 + 1, j0 + 1);
	}
	wt[0] = (1 - fx) * (1 - fy);
	wt[1] = fx * (1 - fy);
	wt[2] = (1 - fx) * fy;
	wt[3] = fx * fy;
}

double cs_sample_scalar(const CubeSphere *cs, const float *fld, vec3 p)
{
	int idx[4];
	double wt[4];
	cs_taps(cs, p, idx, wt);
	return wt[0] * fld[idx[0]] + wt[1] * fld[idx[1]] + wt[2] * fld[idx[2]] + wt[3] * fld[idx[3]];
}

vec3 cs_sample_vec(const CubeSphere *cs, const float *fld3, vec3 p)
{
	int idx[4];
	double wt[4];
	cs_taps(cs, p, idx, wt);
	vec3 acc = v3(0, 0, 0);
	for (int k = 0; k < 4; k++)
		acc = v3_add(acc, v3_scale(v3_load(fld3 + 3 * idx[k]), wt[k]));
	return v3_tangent(acc, p);
}

/* ---------------- Voronoi construction ---------------- */

#define POLY_MAX 64
#define LINK_MAX 24

typedef struct {
	double a, b;
	int edge; /* candidate whose bisector carries the edge starting here; -1 = box side */
} PolyV;

typedef struct {
	int cell;
	double len;
} VLink;

/* Keep the part of polygon `in` where A a + B b + C >= 0. */
static int halfplane_clip(const PolyV *in, int n, PolyV *out, double A, double B, double C, int edge)
{
	int m = 0;
	for (int k = 0; k < n; k++) {
		const PolyV *s = &in[k], *e = &in[(k + 1) % n];
		double fs = A * s->a + B * s->b + C;
		double fe = A * e->a + B * e->b + C;
		int s_in = fs >= 0, e_in = fe >= 0;
		if (s_in && m < POLY_MAX)
			out[m++] = *s;
		if (s_in != e_in && m < POLY_MAX) {
			double t = fs / (fs - fe);
			PolyV x;
			x.a = s->a + t * (e->a - s->a);
			x.b = s->b + t * (e->b - s->b);
			/* leaving: the new edge runs along the clip line; entering: the rest of
			 * the original edge keeps its label */
			x.edge = s_in ? edge : s->edge;
			out[m++] = x;
		}
	}
	return m;
}

static double tri_area(vec3 a, vec3 b, vec3 c)
{
	double num = fabs(v3_dot(a, v3_cross(b, c)));
	double den = 1.0 + v3_dot(a, b) + v3_dot(b, c) + v3_dot(c, a);
	return 2.0 * atan2(num, den);
}

/* Voronoi links (neighbour, edge length) and area of cell c. */
static int voronoi_links(const CubeSphere *cs, int c, VLink *links, double *area, int *open)
{
	int N = cs->N;
	int f = c / (N * N), i = (c % (N * N)) % N, j = (c % (N * N)) / N;
	vec3 p = cs_pos(cs, c);
	vec3 e1 = v3_normalize(v3_tangent(cs_face_right[f], p));
	vec3 e2 = v3_cross(p, e1);

	int cand[LINK_MAX], nc = 0;
	for (int dj = -2; dj <= 2; dj++)
		for (int di = -2; di <= 2; di++) {
			if (di == 0 && dj == 0)
				continue;
			int q = resolve_cell(cs, f, i + di, j + dj);
			int seen = q == c;
			for (int k = 0; k < nc && !seen; k++)
				seen = cand[k] == q;
			if (!seen)
				cand[nc++] = q;
		}

	PolyV buf[2][POLY_MAX];
	double S = 4 * cs->h;
	int n = 4, cur = 0;
	buf[0][0] = (PolyV) { -S, -S, -1 };
	buf[0][1] = (PolyV) { S, -S, -1 };
	buf[0][2] = (PolyV) { S, S, -1 };
	buf[0][3] = (PolyV) { -S, S, -1 };
	for (int k = 0; k < nc; k++) {
		vec3 m = v3_sub(p, cs_pos(cs, cand[k]));
		n = halfplane_clip(buf[cur], n, buf[cur ^ 1], v3_dot(e1, m), v3_dot(e2, m), v3_dot(p, m), cand[k]);
		cur ^= 1;
	}

	const PolyV *poly = buf[cur];
	int nl = 0;
	double ar = 0;
	*open = 0;
	for (int k = 0; k < n; k++) {
		const PolyV *s = &poly[k], *e = &poly[(k + 1) % n];
		vec3 V0 = v3_normalize(v3_add(p, v3_add(v3_scale(e1, s->a), v3_scale(e2, s->b))));
		vec3 V1 = v3_normalize(v3_add(p, v3_add(v3_scale(e1, e->a), v3_scale(e2, e->b))));
		ar += tri_area(p, V0, V1);
		if (s->edge < 0) {
			*open = 1;
			continue;
		}
		double len = v3_angle(V0, V1);
		int l = 0;
		while (l < nl && links[l].cell != s->edge)
			l++;
		if (l < nl) {
			links[l].len += len;
		} else if (nl < LINK_MAX) {
			links[nl].cell = s->edge;
			links[nl].len = len;
			nl++;
		}
	}
	*area = ar;
	return nl;
}

/* stable sort, longest edge first */
static void sort_links(VLink *v, int n)
{
	for (int a = 1; a < n; a++) {
		VLink t = v[a];
		int b = a;
		while (b > 0 && v[b - 1].len < t.len) {
			v[b] = v[b - 1];
			b--;
		}
		v[b] = t;
	}
}

void cs_free(CubeSphere *cs)
{
	ggb_grid_dev_release(cs);
	free(cs->pos);
	free(cs->area);
	free(cs->nnb);
	free(cs->nbr);
	free(cs->w);
	free(cs->gc);
	free(cs->elen);
	free(cs->ghost);
	memset(cs, 0, sizeof(*cs));
}

int cs_init(CubeSphere *cs, int N)
{
	memset(cs, 0, sizeof(*cs));
	if (N < 1)
		return -1;
	cs->N = N;
	cs->ncells = 6 * N * N;
	cs->h = (M_PI / 2) / N;
	int n = cs->ncells;
	size_t nk = (size_t) CS_MAXNB * n;
	int P = N + 2;
	cs->pos = malloc(sizeof(float) * 3 * (size_t) n);
	cs->area = malloc(sizeof(float) * (size_t) n);
	cs->nnb = calloc((size_t) n, 1);
	cs->nbr = calloc(nk, sizeof(int));
	cs->w = calloc(nk, sizeof(float));
	cs->elen = calloc(nk, sizeof(float));
	cs->gc = calloc(3 * nk, sizeof(float));
	cs->ghost = malloc(sizeof(int) * 6 * (size_t) P * P);
	float *len1 = malloc(sizeof(float) * nk);
	float *len2 = malloc(sizeof(float) * nk);
	double *carea = malloc(sizeof(double) * (size_t) n);
	int *copen = malloc(sizeof(int) * (size_t) n);
	int *casym = malloc(sizeof(int) * (size_t) n);
	if (!cs->pos || !cs->area || !cs->nnb || !cs->nbr || !cs->w || !cs->elen || !cs->gc ||
	    !cs->ghost || !len1 || !len2 || !carea || !copen || !casym) {
		fprintf(stderr, "cs_init: out of memory\n");
		free(len1); free(len2); free(carea); free(copen); free(casym);
		cs_free(cs);
		return -1;
	}
	double h = cs->h;

	#pragma omp parallel for schedule(static)
	for (int c = 0; c < n; c++) {
		int f = c / (N * N), rem = c % (N * N);
		v3_store(cs->pos + 3 * c, centre_dir(cs, f, rem % N, rem / N));
	}

	/* Voronoi links, longest first, negligible slivers dropped */
	double minlen = 1e-4 * h;
	#pragma omp parallel for schedule(dynamic, 256)
	for (int c = 0; c < n; c++) {
		VLink links[LINK_MAX];
		int op;
		int nl = voronoi_links(cs, c, links, &carea[c], &op);
		copen[c] = op;
		cs->area[c] = (float) carea[c];
		sort_links(links, nl);
		int m = 0;
		while (m < nl && m < CS_MAXNB && links[m].len >= minlen) {
			cs->nbr[CS_MAXNB * c + m] = links[m].cell;
			len1[CS_MAXNB * c + m] = (float) links[m].len;
			m++;
		}
		cs->nnb[c] = (unsigned char) m;
	}
	double total = 0;
	int open = 0;
	for (int c = 0; c < n; c++) {
		total += carea[c];
		open += copen[c];
	}
	cs->total_area = total;
	cs->open_cells = open;

	/* reciprocity: mutual links keep the mean of both edge lengths */
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < n; c++) {
		casym[c] = 0;
		for (int k = 0; k < cs->nnb[c]; k++) {
			int q = cs->nbr[CS_MAXNB * c + k];
			int back = -1;
			for (int kk = 0; kk < cs->nnb[q]; kk++)
				if (cs->nbr[CS_MAXNB * q + kk] == c)
					back = kk;
			double l = len1[CS_MAXNB * c + k];
			if (back < 0) {
				if (l > 1e-2 * h)
					casym[c]++;
				len2[CS_MAXNB * c + k] = -1.0f;
			} else {
				len2[CS_MAXNB * c + k] = (float) (0.5 * (l + len1[CS_MAXNB * q + back]));
			}
		}
	}
	int asym = 0;
	for (int c = 0; c < n; c++)
		asym += casym[c];
	cs->asym_count = asym;

	/* final stencil: Laplacian weights and least-squares gradient coefficients */
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < n; c++) {
		vec3 p = cs_pos(cs, c);
		int nb[CS_MAXNB], m = 0;
		double ln[CS_MAXNB];
		for (int k = 0; k < cs->nnb[c]; k++) {
			if (len2[CS_MAXNB * c + k] < 0)
				continue;
			nb[m] = cs->nbr[CS_MAXNB * c + k];
			ln[m] = len2[CS_MAXNB * c + k];
			m++;
		}
		cs->nnb[c] = (unsigned char) m;
		int f = c / (N * N);
		vec3 e1 = v3_normalize(v3_tangent(cs_face_right[f], p));
		vec3 e2 = v3_cross(p, e1);
		double M00 = 0, M01 = 0, M11 = 0;
		double tx[CS_MAXNB], ty[CS_MAXNB], wk[CS_MAXNB];
		for (int k = 0; k < m; k++) {
			vec3 q = cs_pos(cs, nb[k]);
			double d = v3_angle(p, q);
			cs->nbr[CS_MAXNB * c + k] = nb[k];
			cs->w[CS_MAXNB * c + k] = (float) (ln[k] / d);
			cs->elen[CS_MAXNB * c + k] = (float) ln[k];
			vec3 t = v3_tangent(v3_sub(q, p), p);
			tx[k] = v3_dot(t, e1);
			ty[k] = v3_dot(t, e2);
			wk[k] = ln[k] / (d * d * d);
			M00 += wk[k] * tx[k] * tx[k];
			M01 += wk[k] * tx[k] * ty[k];
			M11 += wk[k] * ty[k] * ty[k];
		}
		for (int k = m; k < CS_MAXNB; k++) {
			cs->nbr[CS_MAXNB * c + k] = 0;
			cs->w[CS_MAXNB * c + k] = 0;
			cs->elen[CS_MAXNB * c + k] = 0;
		}
		double det = M00 * M11 - M01 * M01;
		double i00 = M11 / det, i01 = -M01 / det, i11 = M00 / det;
		for (int k = 0; k < m; k++) {
			double c1 = wk[k] * (i00 * tx[k] + i01 * ty[k]);
			double c2 = wk[k] * (i01 * tx[k] + i11 * ty[k]);
			v3_store(cs->gc + 3 * (CS_MAXNB * c + k), v3_add(v3_scale(e1, c1), v3_scale(e2, c2)));
		}
	}

	/* seam ghost table for bilinear taps: every (f, i, j) with i, j in [-1, N] */
	#pragma omp parallel for schedule(static)
	for (int r = 0; r < 6 * P; r++) {
		int f = r / P, j = r % P - 1;
		for (int i = -1; i <= N; i++)
			cs->ghost[((size_t) f * P + (j + 1)) * P + (i + 1)] = resolve_cell(cs, f, i, j);
	}

	free(len1);
	free(len2);
	free(carea);
	free(copen);
	free(casym);
	if (asym)
		fprintf(stderr, "cs_init: warning: %d non-reciprocal Voronoi links\n", asym);
	if (open)
		fprintf(stderr, "cs_init: warning: %d open Voronoi cells\n", open);
	return 0;
}
