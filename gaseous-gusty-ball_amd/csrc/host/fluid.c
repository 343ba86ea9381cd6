/*
 * Fluid solver, host side: owns the device state of a Fluid and sequences the gfx950
 * kernels of one step (reference src/fluid.c).  No per-cell work happens here.
 *
 * Step order (reference fluid.c:424-478): departures -> absolute vorticity ->
 * BFECC advect/advect/source/limited advect -> relax + un-Coriolis -> remove mean ->
 * warm-started MG-PCG -> velocity (+ max speed).  The PCG loop runs on the device
 * with device-resident scalars; the host enqueues iterations ahead (as many as the
 * previous solve needed) and reads the loop flag back once per batch.
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fluid.h"
#include "ggb_ext.h"
#include "ggb_host.h"

#define GGB_VERSION "ggb-amd 0.1 (gfx950)"

/* ---------------- runtime helpers ---------------- */

void ggb_die(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	fprintf(stderr, "ggb: ");
	vfprintf(stderr, fmt, ap);
	fprintf(stderr, "\n");
	va_end(ap);
	exit(1);
}

int ggb_ensure_device(void) { return ggbk_init(); }

/* -1: not yet read from the environment */
static int g_sync_mirrors = -1;

int ggb_sync_mirrors(void)
{
	int v = __atomic_load_n(&g_sync_mirrors, __ATOMIC_RELAXED);
	if (v < 0) {
		const char *e = getenv("GGB_SYNC_MIRRORS");
		v = e && *e && strcmp(e, "0") != 0;
		__atomic_store_n(&g_sync_mirrors, v, __ATOMIC_RELAXED);
	}
	return v;
}

void ggb_set_sync_mirrors(int on) { __atomic_store_n(&g_sync_mirrors, on ? 1 : 0, __ATOMIC_RELAXED); }

int ggb_device_init(void) { return ggbk_init(); }

const char *ggb_version(void) { return GGB_VERSION; }

void ggb_sync(void)
{
	if (ggbk_sync())
		ggb_die("device error: %s", ggbk_last_error());
}

void *ggb_dmalloc(size_t bytes, const char *what)
{
	void *p = ggbk_malloc(bytes);
	if (!p)
		ggb_die("%s: out of device memory (%zu bytes)", what, bytes);
	return p;
}

static void upload(void *dst, const void *src, size_t bytes)
{
	if (ggbk_h2d(dst, src, bytes))
		ggb_die("upload failed: %s", ggbk_last_error());
}

static void download(void *dst, const void *src, size_t bytes)
{
	if (ggbk_d2h(dst, src, bytes))
		ggb_die("download failed: %s", ggbk_last_error());
}

/* ---------------- grid tables on the device ---------------- */

static void *grid_buf(struct ggb_grid_dev *gd, size_t bytes, const void *host)
{
	void *p = ggb_dmalloc(bytes, "grid tables");
	gd->bufs[gd->nbufs++] = p;
	if (host)
		upload(p, host, bytes);
	return p;
}

/* Stencil of the Laplacian passes in compressed form.  The Voronoi weights repeat under
 * the cube's symmetries (cells whose weight tuples, in slot order, are bitwise equal: about
 * 47 cells per tuple), so each distinct tuple is stored once as a class row of 8 floats
 * (zero-padded) that stays in cache.  Per cell: the class, and the neighbours as 4-bit
 * codes (dy + 1) << 2 | (dx + 1) of their index offsets dy N + dx, dy, dx in {-1, 0, 1}
 * (self, code 5, is the padding); a cell with any other offset (across a face seam) has
 * its class tagged in bit 31 and an explicit row of 8 indices (self-padded) instead. */
static int offset_code(int off, int N)
{
	for (int dy = -1; dy <= 1; dy++)
		for (int dx = -1; dx <= 1; dx++)
			if (off == dy * N + dx)
				return (dy + 1) << 2 | (dx + 1);
	return -1;
}

static uint64_t tuple_hash(const uint32_t *v, int m)
{
	uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t) m;
	for (int k = 0; k < m; k++) {
		h ^= v[k];
		h *= 0xff51afd7ed558ccdull;
		h ^= h >> 32;
	}
	return h;
}

int ggb_stencil_compress(const CubeSphere *cs, int *rec, float *cw, int *xn, int *ncls_out, int *nirr_out)
{
	int n = cs->ncells, N = cs->N;
	size_t hcap = 16;
	while (hcap < 2 * (size_t) n)
		hcap <<= 1;
	int *htab = malloc(sizeof(int) * hcap);
	if (!htab)
		return -1;
	memset(htab, 0xff, sizeof(int) * hcap);
	int ncls = 0, nirr = 0;
	for (int i = 0; i < n; i++) {
		uint32_t key[9];
		key[0] = cs->nnb[i];
		for (int k = 0; k < 8; k++) {
			float wv = k < cs->nnb[i] ? cs->w[(size_t) CS_MAXNB * i + k] : 0.0f;
			memcpy(&key[1 + k], &wv, sizeof(float));
		}
		size_t hpos = tuple_hash(key, 9) & (hcap - 1);
		int cls;
		for (;;) {
			cls = htab[hpos];
			if (cls < 0) {
				cls = ncls++;
				memcpy(&cw[8 * (size_t) cls], &key[1], 8 * sizeof(float));
				htab[hpos] = cls;
				break;
			}
			if (!memcmp(&cw[8 * (size_t) cls], &key[1], 8 * sizeof(float)))
				break;
			hpos = (hpos + 1) & (hcap - 1);
		}
		unsigned code = 0;
		int regular = 1;
		for (int k = cs->nnb[i]; k < 8; k++)
			code |= 5u << (4 * k);
		for (int k = 0; k < cs->nnb[i] && regular; k++) {
			int c = offset_code(cs->nbr[(size_t) CS_MAXNB * i + k] - i, N);
			if (c < 0)
				regular = 0;
			else
				code |= (unsigned) c << (4 * k);
		}
		if (regular) {
			rec[2 * (size_t) i] = cls;
			rec[2 * (size_t) i + 1] = (int) code;
		} else {
			for (int k = 0; k < 8; k++)
				xn[8 * (size_t) nirr + k] = k < cs->nnb[i] ? cs->nbr[(size_t) CS_MAXNB * i + k] : i;
			rec[2 * (size_t) i] = (int) (0x80000000u | (unsigned) cls);
			rec[2 * (size_t) i + 1] = nirr++;
		}
	}
	free(htab);
	*ncls_out = ncls;
	*nirr_out = nirr;
	return 0;
}

static void build_stencil_classes(struct ggb_grid_dev *gd, ggbk_grid *g, const CubeSphere *cs)
{
	int n = cs->ncells;
	float *cw = malloc(sizeof(float) * 8 * (size_t) n); /* class rows (at most n) */
	int *rec = malloc(sizeof(int) * 2 * (size_t) n);
	int *xn = malloc(sizeof(int) * 8 * (size_t) n);
	int ncls, nirr;
	if (!cw || !rec || !xn || ggb_stencil_compress(cs, rec, cw, xn, &ncls, &nirr))
		ggb_die("out of memory");
	g->srec = grid_buf(gd, sizeof(int) * 2 * (size_t) n, rec);
	g->scw = grid_buf(gd, sizeof(float) * 8 * (size_t) ncls, cw);
	g->sxn = grid_buf(gd, sizeof(int) * 8 * (size_t) (nirr > 0 ? nirr : 1), nirr > 0 ? xn : NULL);
	g->ncls = ncls;
	g->nirr = nirr;
	free(cw);
	free(rec);
	free(xn);
}

ggbk_grid ggb_grid_dev(const CubeSphere *csc)
{
	CubeSphere *cs = (CubeSphere *) csc; /* the device cache is library-managed */
	if (cs->dev)
		return cs->dev->g;
	if (ggb_ensure_device())
		ggb_die("no GPU available");
	int n = cs->ncells, N = cs->N;
	int maxnb = 0;
	for (int i = 0; i < n; i++)
		if (cs->nnb[i] > maxnb)
			maxnb = cs->nnb[i];
	if (maxnb < 1)
		maxnb = 1;
	struct ggb_grid_dev *gd = calloc(1, sizeof(*gd));
	if (!gd)
		ggb_die("out of memory");
	size_t nk = (size_t) maxnb * n;
	int *nbr = malloc(sizeof(int) * nk);
	float *w = malloc(sizeof(float) * nk);
	float *el = malloc(sizeof(float) * nk);
	float *gc = malloc(sizeof(float) * 3 * nk);
	float *dinv = malloc(sizeof(float) * (size_t) n);
	if (!nbr || !w || !el || !gc || !dinv)
		ggb_die("out of memory");
	for (int i = 0; i < n; i++) {
		double s = 0;
		for (int k = 0; k < maxnb; k++) {
			int in = k < cs->nnb[i];
			size_t src = (size_t) CS_MAXNB * i + k;
			nbr[(size_t) k * n + i] = in ? cs->nbr[src] : i;
			w[(size_t) k * n + i] = in ? cs->w[src] : 0.0f;
			el[(size_t) k * n + i] = in ? cs->elen[src] : 0.0f;
			for (int c = 0; c < 3; c++)
				gc[((size_t) k * 3 + c) * n + i] = in ? cs->gc[3 * src + c] : 0.0f;
		}
		/* dinv = 1 / sum w (reference fluid.c:182-187, 58-63) */
		for (int k = 0; k < cs->nnb[i]; k++)
			s += cs->w[(size_t) CS_MAXNB * i + k];
		dinv[i] = (float) (1.0 / s);
	}
	ggbk_grid g;
	memset(&g, 0, sizeof(g));
	g.N = N;
	g.n = n;
	g.maxnb = maxnb;
	g.h = cs->h;
	g.total_area = cs->total_area;
	g.pos = grid_buf(gd, sizeof(float) * 3 * (size_t) n, cs->pos);
	g.area = grid_buf(gd, sizeof(float) * (size_t) n, cs->area);
	g.nnb = grid_buf(gd, (size_t) n, cs->nnb);
	g.nbr = grid_buf(gd, sizeof(int) * nk, nbr);
	g.w = grid_buf(gd, sizeof(float) * nk, w);
	g.elen = grid_buf(gd, sizeof(float) * nk, el);
	g.gc = grid_buf(gd, sizeof(float) * 3 * nk, gc);
	g.ghost = grid_buf(gd, sizeof(int) * 6 * (size_t) (N + 2) * (N + 2), cs->ghost);
	g.dinv = grid_buf(gd, sizeof(float) * (size_t) n, dinv);
	build_stencil_classes(gd, &g, cs);
	free(nbr);
	free(w);
	free(el);
	free(gc);
	free(dinv);
	gd->g = g;
	cs->dev = gd;
	return g;
}

void ggb_grid_dev_release(CubeSphere *cs)
{
	if (!cs->dev)
		return;
	ggb_sync();
	for (int k = 0; k < cs->dev->nbufs; k++)
		ggbk_free(cs->dev->bufs[k]);
	free(cs->dev);
	cs->dev = NULL;
}

/* ---------------- fluid device state ---------------- */

/* host copy of one multigrid level, kept for the levels whose V-cycle is
 * precomputed as a dense matrix */
typedef struct {
	const CubeSphere *cs;
	float *dinv;
	int *ptap;   /* [nf][4] taps into the next coarser level */
	float *pw;
	int *rstart, *ridx;
	float *rw;
} HostLevel;

struct ggb_fluid_dev {
	int prec;
	int T, TM, TZ;          /* state, multigrid and preconditioned-residual types */
	size_t tsz, tmsz, tzsz;
	int n;
	ggbk_grid g0;
	int nlev;
	int dense_level;        /* level whose sub-V-cycle is a dense matrix, or -1 */
	int cheb;               /* mixed precision: the tiled Chebyshev V-cycle (ggbk_vcycle_cheb) */
	/* mixed precision: the solve of a step starts from psi extrapolated in time */
	double *psi_prev[2];    /* psi of the previous two steps (FP64, [0] the newer), or NULL */
	int prev_valid;         /* how many of them (and of dt_prev) hold past steps: 0..2 */
	double dt_prev[2];      /* the steps that led to them: [0] to psi_prev[0] */
	int no_extrap;          /* set while the adjoint replays steps */
	CubeSphere *coarse;     /* [nlev]; [0] unused */
	ggbk_level *levels;     /* [nlev] */
	HostLevel *hl;          /* [nlev] */
	void **xbufs;
	int nxbufs;
	void *omega, *omega_band, *psi, *psi_alt, *vel, *q0, *q1, *q2, *dep_b, *dep_f, *rbuf[2], *pbuf[2], *Ap, *z;
	double *Az;             /* mixed-precision PCG: A z of the last V-cycle */
	ggbk_scalars *s;        /* device */
	ggbk_scalars *hs;       /* pinned host copy */
	ggbk_red red;
	int nb0;                /* blocks of a level-0 pass */
	int predict;            /* iterations the previous solve needed */
	int fix_its;            /* >= 0: the next solves run exactly this many iterations */
	int vmax_valid;         /* s->vmax2 matches vel */
	unsigned vmax_seq;      /* max-speed folds enqueued */
	unsigned vmax_read;     /* vmax_seq at the last read of the scalars (hs->vmax2 is current
	                         * when they are equal) */
	/* last-synced copies of the host mirrors (detects caller writes) */
	float *shadow_omega, *shadow_band, *shadow_psi, *shadow_vel;
};

#define MAX_LIVE 64
static const Fluid *g_live[MAX_LIVE];
static pthread_mutex_t g_live_mu = PTHREAD_MUTEX_INITIALIZER;

static void live_add(const Fluid *F)
{
	pthread_mutex_lock(&g_live_mu);
	for (int k = 0; k < MAX_LIVE; k++)
		if (!g_live[k]) {
			g_live[k] = F;
			break;
		}
	pthread_mutex_unlock(&g_live_mu);
}

static void live_remove(const Fluid *F)
{
	pthread_mutex_lock(&g_live_mu);
	for (int k = 0; k < MAX_LIVE; k++)
		if (g_live[k] == F)
			g_live[k] = NULL;
	pthread_mutex_unlock(&g_live_mu);
}

const Fluid *ggb_fluid_of_vel(const float *vel)
{
	const Fluid *r = NULL;
	pthread_mutex_lock(&g_live_mu);
	for (int k = 0; k < MAX_LIVE; k++)
		if (g_live[k] && g_live[k]->vel == vel)
			r = g_live[k];
	pthread_mutex_unlock(&g_live_mu);
	return r;
}

const void *ggb_fluid_dev_vel(const Fluid *F, int *T)
{
	*T = F->dev->T;
	return F->dev->vel;
}

static void *xbuf(struct ggb_fluid_dev *D, size_t bytes, const char *what)
{
	void *p = ggb_dmalloc(bytes, what);
	D->xbufs[D->nxbufs++] = p;
	return p;
}

static float *falloc(size_t n)
{
	float *f = calloc(n ? n : 1, sizeof(float));
	if (!f)
		ggb_die("fluid: out of memory");
	return f;
}

/* reference fluid.c:44-47 */
static int mg_levels(int N)
{
	int nlev = 1;
	while ((N >> nlev) >= 4 && ((N >> (nlev - 1)) % 2) == 0)
		nlev++;
	return nlev;
}

/* Whether the transfer's indices follow the face-interior pattern the device passes
 * compute instead of reading: with Nf = 2 Nc, fine cell (f, x, y) with 1 <= x, y <= Nf - 2
 * has its taps at coarse (f, (x - 1) >> 1, (y - 1) >> 1) and the three cells right, up and
 * up-right of it; coarse cell (f, I, J) with 1 <= I, J <= Nc - 2 has exactly the 16 fine
 * cells x in [2I - 1, 2I + 2], y in [2J - 1, 2J + 2] in its row of R, in fine-index order,
 * fine cell (x, y) with the weight of its tap 2 [y <= 2J] + [x <= 2I].  Checked cell by
 * cell on the tables built above; the weights stay tabulated. */
static int transfer_is_analytic(const CubeSphere *fine, const CubeSphere *coarse, const int *ptap,
				const int *rstart, const int *ridx, const float *rw, const float *pw)
{
	int Nf = fine->N, Nc = coarse->N;
	if (Nf != 2 * Nc || Nc < 3)
		return 0;
	for (int f = 0; f < 6; f++)
		for (int y = 1; y <= Nf - 2; y++)
			for (int x = 1; x <= Nf - 2; x++) {
				size_t i = ((size_t) f * Nf + y) * Nf + x;
				int b = (f * Nc + ((y - 1) >> 1)) * Nc + ((x - 1) >> 1);
				const int *t = ptap + 4 * i;
				if (t[0] != b || t[1] != b + 1 || t[2] != b + Nc || t[3] != b + Nc + 1)
					return 0;
			}
	for (int f = 0; f < 6; f++)
		for (int J = 1; J <= Nc - 2; J++)
			for (int I = 1; I <= Nc - 2; I++) {
				int c = (f * Nc + J) * Nc + I;
				if (rstart[c + 1] - rstart[c] != 16)
					return 0;
				for (int e = 0; e < 16; e++) {
					int y = 2 * J - 1 + e / 4, x = 2 * I - 1 + e % 4;
					int i = (f * Nf + y) * Nf + x;
					int k = 2 * (y <= 2 * J) + (x <= 2 * I);
					if (ridx[rstart[c] + e] != i || memcmp(&rw[rstart[c] + e], &pw[4 * (size_t) i + k], 4))
						return 0;
				}
			}
	return 1;
}

/* Bilinear transfer from the coarse grid to the fine grid: taps of every fine centre on
 * the coarse grid, and R = P^T as CSR with fine indices ascending (fluid.c:64-96).  The
 * host tables are left in *hl.  With exact, the face-interior cells (whose taps follow the
 * bilinear pattern) take the exact bilinear weights in equiangular coordinates, products
 * of 1/4 and 3/4 by the cell's parity, instead of those evaluated at the float-rounded
 * cell centres (which differ by up to ~1e-5): the mixed-precision preconditioner's own
 * transfer, whose device passes compute these weights instead of reading them. */
static ggbk_xfer build_transfer(struct ggb_fluid_dev *D, const CubeSphere *fine, const CubeSphere *coarse,
				HostLevel *hl, int exact)
{
	int nf = fine->ncells, nc = coarse->ncells;
	int *ptap = malloc(sizeof(int) * 4 * (size_t) nf);
	float *pw = malloc(sizeof(float) * 4 * (size_t) nf);
	int *rstart = calloc((size_t) nc + 1, sizeof(int));
	int *ridx = malloc(sizeof(int) * 4 * (size_t) nf);
	float *rw = malloc(sizeof(float) * 4 * (size_t) nf);
	int *cur = malloc(sizeof(int) * (size_t) nc);
	if (!ptap || !pw || !rstart || !ridx || !rw || !cur)
		ggb_die("fluid: out of memory");
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < nf; i++) {
		int idx[4];
		double wt[4];
		cs_taps(coarse, cs_pos(fine, i), idx, wt);
		for (int k = 0; k < 4; k++) {
			ptap[4 * i + k] = idx[k];
			pw[4 * i + k] = (float) wt[k];
		}
	}
	if (exact && fine->N == 2 * coarse->N) {
		int Nf = fine->N, Nc = coarse->N;
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < nf; i++) {
			int f = i / (Nf * Nf), y = (i / Nf) % Nf, x = i % Nf;
			if (x < 1 || x > Nf - 2 || y < 1 || y > Nf - 2)
				continue;
			int b = (f * Nc + ((y - 1) >> 1)) * Nc + ((x - 1) >> 1);
			const int *t = ptap + 4 * (size_t) i;
			if (t[0] != b || t[1] != b + 1 || t[2] != b + Nc || t[3] != b + Nc + 1)
				continue;
			double ax = (x & 1) ? 0.75 : 0.25, ay = (y & 1) ? 0.75 : 0.25;
			pw[4 * (size_t) i + 0] = (float) (ax * ay);
			pw[4 * (size_t) i + 1] = (float) ((1.0 - ax) * ay);
			pw[4 * (size_t) i + 2] = (float) (ax * (1.0 - ay));
			pw[4 * (size_t) i + 3] = (float) ((1.0 - ax) * (1.0 - ay));
		}
	}
	for (int i = 0; i < 4 * nf; i++)
		rstart[ptap[i] + 1]++;
	for (int c = 0; c < nc; c++)
		rstart[c + 1] += rstart[c];
	memcpy(cur, rstart, sizeof(int) * (size_t) nc);
	for (int i = 0; i < nf; i++)
		for (int k = 0; k < 4; k++) {
			int c = ptap[4 * i + k];
			ridx[cur[c]] = i;
			rw[cur[c]] = pw[4 * i + k];
			cur[c]++;
		}
	/* R as padded rows (ELL, column-major), fine index ascending within a row */
	int rmax = 0;
	for (int c = 0; c < nc; c++)
		if (rstart[c + 1] - rstart[c] > rmax)
			rmax = rstart[c + 1] - rstart[c];
	if (rmax > GGBK_RMAX)
		ggb_die("fluid: restriction row of %d entries exceeds %d", rmax, GGBK_RMAX);
	int rpad = rmax <= 18 ? 18 : GGBK_RMAX; /* the restriction kernel is built for these widths */
	int *rr = calloc((size_t) 2 * rpad * nc, sizeof(int)); /* padding: fine cell 0, weight 0 */
	int *pt = malloc(sizeof(int) * 8 * (size_t) nf);
	if (!rr || !pt)
		ggb_die("fluid: out of memory");
	for (int c = 0; c < nc; c++)
		for (int k = rstart[c]; k < rstart[c + 1]; k++) {
			int e = k - rstart[c];
			size_t o = ((size_t) (e / 2) * nc + c) * 4 + 2 * (e % 2);
			rr[o] = ridx[k];
			memcpy(&rr[o + 1], &rw[k], sizeof(float));
		}
	for (size_t f = 0; f < (size_t) nf; f++)
		for (int k = 0; k < 4; k++) {
			pt[8 * f + 2 * k] = ptap[4 * f + k];
			memcpy(&pt[8 * f + 2 * k + 1], &pw[4 * f + k], sizeof(float));
		}
	ggbk_xfer x;
	memset(&x, 0, sizeof(x));
	x.nf = nf;
	x.nc = nc;
	x.rmax = rmax;
	x.rpad = rpad;
	x.Nc = coarse->N;
	/* computed indices pay off where the tables stream from HBM; the small levels' tables
	 * stay in L2, and there the index arithmetic costs more than the loads it saves */
	x.rana = transfer_is_analytic(fine, coarse, ptap, rstart, ridx, rw, pw);
	/* with exact weights nothing is read for the face-interior cells: computed at every size */
	x.exact = exact && x.rana;
	x.analytic = x.rana && (x.exact || nc >= 65536);
	void *d;
	d = xbuf(D, sizeof(int) * 8 * (size_t) nf, "transfer");
	upload(d, pt, sizeof(int) * 8 * (size_t) nf);
	x.pt = d;
	if (!x.exact) {
		d = xbuf(D, sizeof(float) * 4 * (size_t) nf, "transfer");
		upload(d, pw, sizeof(float) * 4 * (size_t) nf);
		x.pw = d;
	}
	if (x.rana && !x.exact) {
		/* the weights of the face-interior rows alone, coalesced per entry group */
		float *w16 = calloc(16 * (size_t) nc, sizeof(float));
		if (!w16)
			ggb_die("fluid: out of memory");
		int Nc = coarse->N;
		for (int c = 0; c < nc; c++) {
			int J = (c / Nc) % Nc, I = c % Nc;
			if (I < 1 || I > Nc - 2 || J < 1 || J > Nc - 2)
				continue;
			for (int e = 0; e < 16; e++)
				w16[((size_t) (e / 4) * nc + c) * 4 + e % 4] = rw[rstart[c] + e];
		}
		d = xbuf(D, sizeof(float) * 16 * (size_t) nc, "transfer");
		upload(d, w16, sizeof(float) * 16 * (size_t) nc);
		x.rw16 = d;
		free(w16);
	}
	d = xbuf(D, sizeof(int) * 2 * (size_t) rpad * nc, "transfer");
	upload(d, rr, sizeof(int) * 2 * (size_t) rpad * nc);
	x.rr = d;
	free(rr);
	free(pt);
	free(cur);
	hl->ptap = ptap;
	hl->pw = pw;
	hl->rstart = rstart;
	hl->ridx = ridx;
	hl->rw = rw;
	return x;
}

static void host_level_free(HostLevel *hl)
{
	free(hl->dinv);
	free(hl->ptap);
	free(hl->pw);
	free(hl->rstart);
	free(hl->ridx);
	free(hl->rw);
	memset(hl, 0, sizeof(*hl));
}

/* ---- host V-cycle in double (reference fluid.c:100-146), used only to tabulate the
 * dense operator of the bottom levels ---- */

static double hK(const CubeSphere *cs, const double *x, int i)
{
	double xi = x[i], s = 0;
	for (int k = 0; k < cs->nnb[i]; k++)
		s += cs->w[CS_MAXNB * i + k] * (xi - x[cs->nbr[CS_MAXNB * i + k]]);
	return s;
}

static void h_smooth(const HostLevel *L, const double *b, double *x, double *r, int sweeps)
{
	int n = L->cs->ncells;
	for (int s = 0; s < sweeps; s++) {
		for (int i = 0; i < n; i++)
			r[i] = b[i] - hK(L->cs, x, i);
		for (int i = 0; i < n; i++)
			x[i] += 0.8 * L->dinv[i] * r[i];
	}
}

/* deg Chebyshev steps for D^-1 K on [1/2, 2] from x (the mixed-precision smoother of
 * mgt_k.hip, whose first step is the damped Jacobi step); d is scratch */
static void h_cheb(const HostLevel *L, const double *b, double *x, double *r, double *d, int deg)
{
	int n = L->cs->ncells;
	double sigma = 5.0 / 3.0, delta = 0.75, rho = 1.0 / sigma;
	for (int k = 0; k < deg; k++) {
		double cd = 0.0, cr = 0.8;
		if (k > 0) {
			double rn = 1.0 / (2 * sigma - rho);
			cd = rn * rho;
			cr = 2 * rn / delta;
			rho = rn;
		}
		for (int i = 0; i < n; i++)
			r[i] = b[i] - hK(L->cs, x, i);
		for (int i = 0; i < n; i++) {
			d[i] = (k ? cd * d[i] : 0.0) + cr * L->dinv[i] * r[i];
			x[i] += d[i];
		}
	}
}

/* x = V_l(b); bufs[l] holds 4 scratch arrays of level l.  deg 1: the reference's damped
 * Jacobi V(1,1) cycle; deg > 1: Chebyshev smoothing of that degree (the mixed mode's) */
static void h_vcycle(const HostLevel *hl, int nlev, int l, const double *b, double *x, double **bufs, int deg)
{
	const HostLevel *L = &hl[l];
	int n = L->cs->ncells;
	double *r = bufs[l], *d = bufs[l] + 3 * (size_t) n;
	memset(x, 0, sizeof(double) * n);
	if (l == nlev - 1) {
		h_smooth(L, b, x, r, 60);
		return;
	}
	if (deg > 1)
		h_cheb(L, b, x, r, d, deg);
	else
		h_smooth(L, b, x, r, 1);
	for (int i = 0; i < n; i++)
		r[i] = b[i] - hK(L->cs, x, i);
	const HostLevel *C = &hl[l + 1];
	int nc = C->cs->ncells;
	double *bc = bufs[l + 1] + nc, *xc = bufs[l + 1] + 2 * (size_t) nc;
	for (int c = 0; c < nc; c++) {
		double acc = 0;
		for (int k = L->rstart[c]; k < L->rstart[c + 1]; k++)
			acc += L->rw[k] * r[L->ridx[k]];
		bc[c] = acc;
	}
	h_vcycle(hl, nlev, l + 1, bc, xc, bufs, deg);
	for (int i = 0; i < n; i++) {
		const int *t = L->ptap + 4 * i;
		const float *w = L->pw + 4 * i;
		x[i] += 1.0 * (w[0] * xc[t[0]] + w[1] * xc[t[1]] + w[2] * xc[t[2]] + w[3] * xc[t[3]]);
	}
	if (deg > 1)
		h_cheb(L, b, x, r, d, deg);
	else
		h_smooth(L, b, x, r, 1);
}

/* the same on the Galerkin hierarchy of the mixed precision (levels l.. of gal, the
 * transfers of hl): Chebyshev smoothing of degree deg, 60 damped Jacobi sweeps on the
 * coarsest level */
static double gK(const GalOp *G, const double *x, int i)
{
	const int *nb = G->nbr + (size_t) GGBK_WNB * i;
	const double *w = G->w + (size_t) GGBK_WNB * i;
	double xi = x[i], s = 0;
	for (int k = 0; k < GGBK_WNB; k++)
		s += w[k] * (xi - x[nb[k]]);
	return s;
}

static void g_cheb(const GalOp *G, const double *b, double *x, double *r, double *d, int deg, int jacobi)
{
	int n = G->n;
	double sigma = 5.0 / 3.0, delta = 0.75, rho = 1.0 / sigma;
	for (int k = 0; k < deg; k++) {
		double cd = 0.0, cr = 0.8;
		if (k > 0 && !jacobi) {
			double rn = 1.0 / (2 * sigma - rho);
			cd = rn * rho;
			cr = 2 * rn / delta;
			rho = rn;
		}
		for (int i = 0; i < n; i++)
			r[i] = b[i] - gK(G, x, i);
		for (int i = 0; i < n; i++) {
			d[i] = (k ? cd * d[i] : 0.0) + cr * (double) G->dinv[i] * r[i];
			x[i] += d[i];
		}
	}
}

static void g_vcycle(const HostLevel *hl, const GalOp *gal, int nlev, int l, const double *b, double *x,
		     double **bufs, int deg)
{
	const GalOp *G = &gal[l];
	const HostLevel *L = &hl[l];
	int n = G->n;
	double *r = bufs[l], *d = bufs[l] + 3 * (size_t) n;
	memset(x, 0, sizeof(double) * n);
	if (l == nlev - 1) {
		g_cheb(G, b, x, r, d, 60, 1);
		return;
	}
	g_cheb(G, b, x, r, d, deg, 0);
	for (int i = 0; i < n; i++)
		r[i] = b[i] - gK(G, x, i);
	int nc = gal[l + 1].n;
	double *bc = bufs[l + 1] + nc, *xc = bufs[l + 1] + 2 * (size_t) nc;
	for (int c = 0; c < nc; c++) {
		double acc = 0;
		for (int k = L->rstart[c]; k < L->rstart[c + 1]; k++)
			acc += L->rw[k] * r[L->ridx[k]];
		bc[c] = acc;
	}
	g_vcycle(hl, gal, nlev, l + 1, bc, xc, bufs, deg);
	for (int i = 0; i < n; i++) {
		const int *t = L->ptap + 4 * i;
		const float *w = L->pw + 4 * i;
		x[i] += w[0] * xc[t[0]] + w[1] * xc[t[1]] + w[2] * xc[t[2]] + w[3] * xc[t[3]];
	}
	g_cheb(G, b, x, r, d, deg, 0);
}

#define DENSE_MAX 1536

/* Tabulate the sub-V-cycle of level L as M[i][j] = (V_L e_j)_i and upload it: the
 * reference's V(1,1) cycle (deg 1), or the mixed precision's Chebyshev cycle (deg 3, as the
 * levels above it smooth), or with gal the Chebyshev cycle of the Galerkin hierarchy. */
static void build_dense(struct ggb_fluid_dev *D, int L, int deg, const GalOp *gal)
{
	int n = D->hl[L].cs->ncells;
	double *M = malloc(sizeof(double) * (size_t) n * n);
	if (!M)
		ggb_die("fluid: out of memory");
	#pragma omp parallel
	{
		double **bufs = calloc((size_t) D->nlev, sizeof(double *));
		for (int l = L; l < D->nlev; l++)
			bufs[l] = calloc(4 * (size_t) D->hl[l].cs->ncells, sizeof(double));
		double *e = calloc((size_t) n, sizeof(double)), *x = calloc((size_t) n, sizeof(double));
		#pragma omp for schedule(dynamic, 8)
		for (int j = 0; j < n; j++) {
			memset(e, 0, sizeof(double) * n);
			e[j] = 1.0;
			if (gal)
				g_vcycle(D->hl, gal, D->nlev, L, e, x, bufs, deg);
			else
				h_vcycle(D->hl, D->nlev, L, e, x, bufs, deg);
			for (int i = 0; i < n; i++)
				M[(size_t) i * n + j] = x[i];
		}
		for (int l = L; l < D->nlev; l++)
			free(bufs[l]);
		free(bufs);
		free(e);
		free(x);
	}
	size_t cnt = (size_t) n * n;
	void *d;
	if (D->TM == GGBK_F64) {
		d = xbuf(D, 8 * cnt, "multigrid dense");
		upload(d, M, 8 * cnt);
	} else {
		float *Mf = malloc(4 * cnt);
		if (!Mf)
			ggb_die("fluid: out of memory");
		for (size_t k = 0; k < cnt; k++)
			Mf[k] = (float) M[k];
		d = xbuf(D, 4 * cnt, "multigrid dense");
		upload(d, Mf, 4 * cnt);
		free(Mf);
	}
	free(M);
	D->levels[L].M = d;
	D->dense_level = L;
}

/* ---- face tiles of a level for the fused Chebyshev passes (mgt_k.hip) ---- */

#define MGT_TILE 16

static int cell_is_regular(const CubeSphere *cs, int i)
{
	for (int k = 0; k < cs->nnb[i]; k++)
		if (offset_code(cs->nbr[(size_t) CS_MAXNB * i + k] - i, cs->N) < 0)
			return 0;
	return 1;
}

/* Tiles of 16x16 cells with `deg` rings.  A tile whose rings lie on its face with regular
 * stencils on rings < deg is interior (no tables); any other gets graph tables: its ext
 * cells ring by ring in BFS order (ring r = the unvisited stencil neighbours of ring r - 1,
 * in discovery order) and each stencil neighbour of rings < deg as a slot of that list.
 * With gal the stencil is that Galerkin operator's (two cells wide: an interior tile's
 * square then has 2 deg rings, one more with the restriction, and its cells on rings
 * <= hs - 2 must be regular).  Returns -1 when a graph tile would exceed capmax ext cells. */
static int build_tiles(struct ggb_fluid_dev *D, const CubeSphere *cs, const GalOp *gal, int deg, const HostLevel *rl,
		       int rana, int capmax, ggbk_tiles *out)
{
	/* rl != NULL: tiles of a down pass, whose restriction through rl's R into the next
	 * level is fused; an interior tile then has one more ring (the rows of R reach one ring
	 * beyond the fine tile) and the bilinear row pattern (rana) */
	int rad = gal ? 2 : 1, S = gal ? GGBK_WNB : 8;
	if (deg + 4 > GGBK_TINFO)
		ggb_die("fluid: %d tile rings exceed the tile record", deg);
	int hs = rad * deg + (rl ? 1 : 0);
	int N = cs->N, n = cs->ncells, tpe = N / MGT_TILE, ntiles = 6 * tpe * tpe;
	unsigned char *reg = malloc((size_t) n);
	int *kind = malloc(sizeof(int) * (size_t) ntiles);
	int *tinfo = calloc((size_t) ntiles * GGBK_TINFO, sizeof(int));
	int *stamp = malloc(sizeof(int) * (size_t) n), *loc = malloc(sizeof(int) * (size_t) n);
	size_t ecap = (size_t) capmax * 64, ne = 0;
	int *gcell = malloc(sizeof(int) * ecap);
	unsigned short *gnbr = malloc(sizeof(unsigned short) * S * ecap);
	int *goff = malloc(sizeof(int) * ((size_t) ntiles + 1));
	int *gring = malloc(sizeof(int) * (size_t) ntiles * (deg + 1));
	size_t rcap = rl ? (size_t) 64 * GGBK_TROW * 64 : 1, nr = 0;
	unsigned short *grow = malloc(sizeof(unsigned short) * rcap);
	float *grw = malloc(sizeof(float) * rcap);
	if (!reg || !kind || !tinfo || !stamp || !loc || !gcell || !gnbr || !goff || !gring || !grow || !grw)
		ggb_die("fluid: out of memory");
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		reg[i] = gal ? gal->reg[i] : (unsigned char) cell_is_regular(cs, i);
		stamp[i] = -1;
	}
	/* GGB_TILES_GRAPH=1: every tile through the graph tables (both forms give the same bits) */
	const char *ge = getenv("GGB_TILES_GRAPH");
	int all_graph = ge && !strcmp(ge, "1");
	int ngraph = 0, cap = 0, rc = 0;
	for (int t = 0; t < ntiles && rc == 0; t++) {
		int f = t / (tpe * tpe), r = t - f * tpe * tpe;
		int J0 = (r / tpe) * MGT_TILE, I0 = (r % tpe) * MGT_TILE;
		int interior = !all_graph && (!rl || rana) && I0 >= hs && J0 >= hs && I0 + MGT_TILE - 1 + hs <= N - 1 &&
			       J0 + MGT_TILE - 1 + hs <= N - 1;
		for (int y = J0 - (hs - rad); interior && y <= J0 + MGT_TILE - 1 + (hs - rad); y++)
			for (int x = I0 - (hs - rad); x <= I0 + MGT_TILE - 1 + (hs - rad); x++)
				if (!reg[(f * N + y) * N + x]) {
					interior = 0;
					break;
				}
		if (interior) {
			kind[t] = -1;
			continue;
		}
		kind[t] = ngraph;
		goff[ngraph] = (int) ne;
		size_t base = ne;
		if (ne + capmax > ecap) {
			ecap = 2 * ecap + capmax;
			gcell = realloc(gcell, sizeof(int) * ecap);
			gnbr = realloc(gnbr, sizeof(unsigned short) * S * ecap);
			if (!gcell || !gnbr)
				ggb_die("fluid: out of memory");
		}
		int cnt = 0;
		for (int y = 0; y < MGT_TILE; y++)
			for (int x = 0; x < MGT_TILE; x++) {
				int c = (f * N + J0 + y) * N + I0 + x;
				stamp[c] = t;
				loc[c] = cnt;
				gcell[base + cnt++] = c;
			}
		if (rl) {
			/* ring 0 also holds the rest of the support of the 8x8 coarse rows */
			int Nc = N / 2;
			if (nr + (size_t) 64 * GGBK_TROW > rcap) {
				rcap = 2 * rcap + (size_t) 64 * GGBK_TROW;
				grow = realloc(grow, sizeof(unsigned short) * rcap);
				grw = realloc(grw, sizeof(float) * rcap);
				if (!grow || !grw)
					ggb_die("fluid: out of memory");
			}
			for (int q = 0; q < 64 && rc == 0; q++) {
				int cc = (f * Nc + J0 / 2 + q / 8) * Nc + I0 / 2 + q % 8;
				int len = rl->rstart[cc + 1] - rl->rstart[cc];
				if (len > GGBK_TROW)
					ggb_die("fluid: a restriction row of %d entries exceeds %d", len, GGBK_TROW);
				for (int k = rl->rstart[cc]; k < rl->rstart[cc + 1]; k++) {
					int j = rl->ridx[k];
					if (stamp[j] == t)
						continue;
					if (cnt == capmax) {
						rc = -1;
						break;
					}
					stamp[j] = t;
					loc[j] = cnt;
					gcell[base + cnt++] = j;
				}
			}
		}
		int *re = gring + (size_t) ngraph * (deg + 1);
		re[0] = cnt;
		int rs = 0;
		for (int ring = 1; ring <= deg && rc == 0; ring++) {
			int rend = cnt;
			for (int e = rs; e < rend && rc == 0; e++) {
				int c = gcell[base + e];
				int nn = gal ? gal->nn[c] : cs->nnb[c];
				for (int k = 0; k < nn; k++) {
					int j = gal ? gal->nbr[(size_t) GGBK_WNB * c + k] : cs->nbr[(size_t) CS_MAXNB * c + k];
					if (stamp[j] == t)
						continue;
					if (cnt == capmax) {
						rc = -1;
						break;
					}
					stamp[j] = t;
					loc[j] = cnt;
					gcell[base + cnt++] = j;
				}
			}
			rs = rend;
			re[ring] = cnt;
		}
		for (int e = 0; e < re[deg - 1] && rc == 0; e++) {
			int c = gcell[base + e];
			int nn = gal ? gal->nn[c] : cs->nnb[c];
			for (int k = 0; k < S; k++) {
				int slot = e;
				if (k < nn) {
					int j = gal ? gal->nbr[(size_t) GGBK_WNB * c + k] : cs->nbr[(size_t) CS_MAXNB * c + k];
					if (stamp[j] != t)
						ggb_die("fluid: tile %d ring cell %d has a neighbour outside its rings", t, c);
					slot = loc[j];
				}
				gnbr[S * (base + e) + k] = (unsigned short) slot;
			}
		}
		for (int e = re[deg - 1]; e < cnt; e++)
			for (int k = 0; k < S; k++)
				gnbr[S * (base + e) + k] = (unsigned short) e;
		if (rl && rc == 0) {
			int Nc = N / 2;
			for (int q = 0; q < 64; q++) {
				int cc = (f * Nc + J0 / 2 + q / 8) * Nc + I0 / 2 + q % 8;
				size_t o = nr + (size_t) q * GGBK_TROW;
				int e = 0;
				for (int k = rl->rstart[cc]; k < rl->rstart[cc + 1]; k++, e++) {
					grow[o + e] = (unsigned short) loc[rl->ridx[k]];
					grw[o + e] = rl->rw[k];
				}
				for (; e < GGBK_TROW; e++) {
					grow[o + e] = 0;
					grw[o + e] = 0.0f;
				}
			}
			nr += (size_t) 64 * GGBK_TROW;
		}
		ne += (size_t) cnt;
		if (cnt > cap)
			cap = cnt;
		ngraph++;
	}
	goff[ngraph] = (int) ne;
	if (rc == 0) {
		/* per tile: kind, first ext cell, ext count, ring ends; whether the interior tiles
		 * are exactly the geometrically interior ones */
		int geo = 1;
		for (int t = 0; t < ntiles; t++) {
			int f = t / (tpe * tpe), r = t - f * tpe * tpe;
			int J0 = (r / tpe) * MGT_TILE, I0 = (r % tpe) * MGT_TILE;
			int gi = I0 >= hs && J0 >= hs && I0 + MGT_TILE - 1 + hs <= N - 1 && J0 + MGT_TILE - 1 + hs <= N - 1;
			int *ti = tinfo + GGBK_TINFO * (size_t) t;
			ti[0] = kind[t];
			if (kind[t] >= 0) {
				int k = kind[t];
				ti[1] = goff[k];
				ti[2] = goff[k + 1] - goff[k];
				for (int q = 0; q <= deg; q++)
					ti[3 + q] = gring[(size_t) k * (deg + 1) + q];
			}
			if (gi != (kind[t] < 0))
				geo = 0;
		}
		/* the weight class of each ext cell, from the device stencil records (a Galerkin
		 * level's weights are per cell: class 0) */
		int *gcc = malloc(sizeof(int) * 2 * (ne + 1));
		int *srec = gal ? NULL : malloc(sizeof(int) * 2 * (size_t) n);
		if (!gcc || (!gal && !srec))
			ggb_die("fluid: out of memory");
		if (!gal) {
			ggbk_grid gg = ggb_grid_dev(cs);
			download(srec, gg.srec, sizeof(int) * 2 * (size_t) n);
		}
		for (size_t e = 0; e < ne; e++) {
			gcc[2 * e] = gcell[e];
			gcc[2 * e + 1] = gal ? 0 : srec[2 * (size_t) gcell[e]] & 0x7fffffff;
		}
		free(srec);
		memset(out, 0, sizeof(*out));
		out->ntiles = ntiles;
		out->deg = deg;
		out->cap = cap;
		out->ngraph = ngraph;
		out->geo = geo;
		out->hs = hs;
		out->nslot = S;
		void *d = xbuf(D, sizeof(int) * GGBK_TINFO * (size_t) ntiles, "multigrid tiles");
		upload(d, tinfo, sizeof(int) * GGBK_TINFO * (size_t) ntiles);
		out->tinfo = d;
		d = xbuf(D, sizeof(int) * 2 * (ne + 1), "multigrid tiles");
		upload(d, gcc, sizeof(int) * 2 * ne);
		out->gcc = d;
		free(gcc);
		d = xbuf(D, sizeof(unsigned short) * S * (ne + 1), "multigrid tiles");
		upload(d, gnbr, sizeof(unsigned short) * S * ne);
		out->gnbr = d;
		if (rl) {
			d = xbuf(D, sizeof(unsigned short) * (nr + 1), "multigrid tiles");
			upload(d, grow, sizeof(unsigned short) * nr);
			out->grow = d;
			d = xbuf(D, sizeof(float) * (nr + 1), "multigrid tiles");
			upload(d, grw, sizeof(float) * nr);
			out->grw = d;
		}
	}
	free(reg);
	free(kind);
	free(tinfo);
	free(stamp);
	free(loc);
	free(gcell);
	free(gnbr);
	free(goff);
	free(gring);
	free(grow);
	free(grw);
	return rc;
}

/* The level whose sub-V-cycle is tabulated as a dense operator: the first coarse level of
 * at most DENSE_MAX cells (none in reference-precision mode, which keeps the reference's
 * float sweeps bit for bit), or -1. */
static int dense_level_of(int prec, int N, int nlev)
{
	if (prec == GGB_PREC_REF32)
		return -1;
	for (int l = 1; l < nlev; l++)
		if (6 * (N >> l) * (N >> l) <= DENSE_MAX)
			return l;
	return -1;
}

/* The mixed-precision preconditioner (Chebyshev smoothing on face tiles, Galerkin coarse
 * levels) needs a dense level below level 0 and every level above it a multiple of 16
 * cells wide; GGB_MIXED_CHEB=0 in the environment selects the reference V(1,1) cycle
 * instead.  setup_cheb builds the tiles of those levels. */
static int cheb_eligible(int prec, int N, int nlev)
{
	if (prec != GGB_PREC_MIXED)
		return 0;
	const char *e = getenv("GGB_MIXED_CHEB");
	if (e && !strcmp(e, "0"))
		return 0;
	int L = dense_level_of(prec, N, nlev);
	if (L < 1)
		return 0;
	for (int l = 0; l < L; l++)
		if ((N >> l) % MGT_TILE)
			return 0;
	return 1;
}

/* capacities of the tiled passes' workgroups (mgt_k.hip): 2 ext cells per thread for
 * level 0 with degree 2, else 3 (level 0 of higher degree, the coarse down passes, the
 * Galerkin levels' passes) */
#define MGT_CAP2 512
#define MGT_CAP3 768

/* The mixed hierarchy's options: GGB_MIXED_GALERKIN=0 keeps the levels' own Voronoi
 * operators (degree 3 below level 0, degree 2 on level 0); GGB_MIXED_DEG0 sets level 0's
 * degree with the Galerkin levels (2-4, default 4: at N = 512 and tol 1e-8, 5 / 4 / 3 CG
 * iterations per step). */
static int mixed_galerkin(void)
{
	const char *e = getenv("GGB_MIXED_GALERKIN");
	return !(e && !strcmp(e, "0"));
}

static int mixed_deg0(void)
{
	const char *e = getenv("GGB_MIXED_DEG0");
	int d = e ? atoi(e) : 4;
	return d < 2 ? 2 : d > 4 ? 4 : d;
}

static void setup_cheb(struct ggb_fluid_dev *D, const CubeSphere *cs, const GalOp *gal, int deg0)
{
	for (int l = 0; l < D->dense_level; l++) {
		const CubeSphere *c = l == 0 ? cs : &D->coarse[l];
		const GalOp *G = gal && l > 0 ? &gal[l] : NULL;
		int deg = l == 0 ? deg0 : G ? 2 : 3;
		int capu = (l == 0 && deg0 == 2) || (l > 0 && !G) ? MGT_CAP2 : MGT_CAP3;
		int capd = l == 0 && deg0 == 2 ? MGT_CAP2 : MGT_CAP3;
		/* level 0's up pass takes one ring more (A z on the tile from z on ring 1) */
		int up_rings = l == 0 ? deg + 1 : deg;
		if (build_tiles(D, c, G, up_rings, NULL, 0, capu, &D->levels[l].tl))
			ggb_die("fluid_init: a multigrid tile of level %d exceeds %d cells", l, capu);
		if (build_tiles(D, c, G, deg, &D->hl[l], D->levels[l].xf.rana, capd, &D->levels[l].td))
			ggb_die("fluid_init: a multigrid tile of level %d exceeds %d cells", l, capd);
		D->levels[l].xpre = xbuf(D, sizeof(float) * (size_t) c->ncells, "multigrid");
		if (G) {
			ggbk_wide *wd = &D->levels[l].wd;
			wd->n = G->n;
			wd->N = G->N;
			void *d = xbuf(D, sizeof(unsigned short) * GGBK_WNB * (size_t) G->n, "multigrid");
			upload(d, G->wh, sizeof(unsigned short) * GGBK_WNB * (size_t) G->n);
			wd->w = d;
			d = xbuf(D, sizeof(float) * (size_t) G->n, "multigrid");
			upload(d, G->dinv, sizeof(float) * (size_t) G->n);
			wd->dinv = d;
		}
	}
	D->Az = xbuf(D, sizeof(double) * (size_t) cs->ncells, "fluid");
}

int fluid_init(Fluid *F, const CubeSphere *cs, const FluidParams *prm)
{
	memset(F, 0, sizeof(*F));
	if (ggb_ensure_device())
		ggb_die("fluid_init: no GPU available (%s)", ggbk_last_error());
	F->cs = cs;
	F->prm = *prm;
	if (F->prm.precision < GGB_PREC_FP64 || F->prm.precision > GGB_PREC_REF32)
		ggb_die("fluid_init: unknown precision %d", F->prm.precision);
	size_t n = cs->ncells;
	F->omega = falloc(n);
	F->omega_band = falloc(n);
	F->psi = falloc(n);
	F->vel = falloc(3 * n);

	struct ggb_fluid_dev *D = calloc(1, sizeof(*D));
	if (!D)
		ggb_die("fluid: out of memory");
	F->dev = D;
	D->prec = F->prm.precision;
	D->T = D->prec == GGB_PREC_REF32 ? GGBK_F32 : GGBK_F64;
	D->TM = D->prec == GGB_PREC_FP64 ? GGBK_F64 : GGBK_F32;
	D->tsz = D->T == GGBK_F64 ? 8 : 4;
	D->tmsz = D->TM == GGBK_F64 ? 8 : 4;
	D->n = (int) n;
	D->g0 = ggb_grid_dev(cs);
	D->nlev = prm->multigrid ? mg_levels(cs->N) : 0;
	D->dense_level = -1;
	D->TZ = prm->multigrid ? D->TM : D->T;
	D->tzsz = D->TZ == GGBK_F64 ? 8 : 4;
	F->nlev = D->nlev;
	D->xbufs = calloc(64 + 8 * (size_t) (D->nlev + 1), sizeof(void *));
	if (!D->xbufs)
		ggb_die("fluid: out of memory");

	size_t ts = D->tsz;
	D->omega = xbuf(D, ts * n, "fluid");
	D->omega_band = xbuf(D, ts * n, "fluid");
	D->psi = xbuf(D, ts * n, "fluid");
	D->psi_alt = xbuf(D, ts * n, "fluid");
	D->vel = xbuf(D, 3 * ts * n, "fluid");
	D->q0 = xbuf(D, ts * n, "fluid");
	D->q1 = xbuf(D, ts * n, "fluid");
	D->q2 = xbuf(D, ts * n, "fluid");
	D->dep_b = xbuf(D, 3 * ts * n, "fluid");
	D->dep_f = xbuf(D, 3 * ts * n, "fluid");
	D->rbuf[0] = xbuf(D, ts * n, "fluid");
	D->rbuf[1] = xbuf(D, ts * n, "fluid");
	D->pbuf[0] = xbuf(D, ts * n, "fluid");
	D->pbuf[1] = xbuf(D, ts * n, "fluid");
	D->Ap = xbuf(D, ts * n, "fluid");
	D->z = xbuf(D, D->tzsz * n, "fluid");
	D->s = xbuf(D, sizeof(ggbk_scalars), "fluid");
	D->hs = ggbk_host_alloc(sizeof(ggbk_scalars));
	if (!D->hs)
		ggb_die("fluid: pinned allocation failed");
	D->nb0 = ggbk_nblocks((int) n);
	D->red.stride = D->nb0;
	D->red.part = xbuf(D, sizeof(double) * GGBK_NSLOTS * (size_t) D->nb0, "fluid");
	D->red.cnt = xbuf(D, sizeof(unsigned) * GGBK_NCNT, "fluid");
	ggbk_zero(D->red.cnt, sizeof(unsigned) * GGBK_NCNT);
	ggbk_zero(D->s, sizeof(ggbk_scalars));
	ggbk_zero(D->omega, ts * n);
	ggbk_zero(D->omega_band, ts * n);
	ggbk_zero(D->psi, ts * n);
	ggbk_zero(D->vel, 3 * ts * n);
	ggbk_zero(D->pbuf[0], ts * n);
	ggbk_zero(D->pbuf[1], ts * n);

	if (D->nlev > 0) {
		D->coarse = calloc((size_t) D->nlev, sizeof(CubeSphere));
		D->levels = calloc((size_t) D->nlev, sizeof(ggbk_level));
		D->hl = calloc((size_t) D->nlev, sizeof(HostLevel));
		if (!D->coarse || !D->levels || !D->hl)
			ggb_die("fluid: out of memory");
		D->levels[0].g = D->g0;
		D->levels[0].r = xbuf(D, D->tmsz * n, "multigrid");
		D->hl[0].cs = cs;
		int cheb = cheb_eligible(D->prec, cs->N, D->nlev);
		for (int l = 1; l < D->nlev; l++) {
			if (cs_init(&D->coarse[l], cs->N >> l))
				ggb_die("fluid_init: coarse grid allocation failed");
			const CubeSphere *fine = l == 1 ? cs : &D->coarse[l - 1];
			D->hl[l].cs = &D->coarse[l];
			D->levels[l].g = ggb_grid_dev(&D->coarse[l]);
			D->levels[l - 1].xf = build_transfer(D, fine, &D->coarse[l], &D->hl[l - 1], cheb);
			size_t nl = D->coarse[l].ncells;
			D->levels[l].x = xbuf(D, D->tmsz * nl, "multigrid");
			D->levels[l].b = xbuf(D, D->tmsz * nl, "multigrid");
			D->levels[l].r = xbuf(D, D->tmsz * nl, "multigrid");
		}
		/* dinv per level on the host (fluid.c:58-63) */
		for (int l = 0; l < D->nlev; l++) {
			const CubeSphere *c = D->hl[l].cs;
			D->hl[l].dinv = malloc(sizeof(float) * (size_t) c->ncells);
			if (!D->hl[l].dinv)
				ggb_die("fluid: out of memory");
			for (int i = 0; i < c->ncells; i++) {
				double sw = 0;
				for (int k = 0; k < c->nnb[i]; k++)
					sw += c->w[CS_MAXNB * i + k];
				D->hl[l].dinv[i] = (float) (1.0 / sw);
			}
		}
		/* the mixed hierarchy's Galerkin operators K_l = R K_(l-1) P, levels 1.. */
		GalOp *gal = NULL;
		if (cheb && mixed_galerkin()) {
			gal = calloc((size_t) D->nlev, sizeof(GalOp));
			if (!gal)
				ggb_die("fluid: out of memory");
			for (int l = 1; l < D->nlev && gal; l++) {
				const HostLevel *h = &D->hl[l - 1];
				if (ggb_galerkin(cs, l == 1 ? NULL : &gal[l - 1], h->ptap, h->pw, h->rstart, h->ridx, h->rw,
						 cs->N >> l, &gal[l])) {
					for (int q = 1; q < l; q++)
						ggb_galop_free(&gal[q]);
					free(gal);
					gal = NULL;
				}
			}
		}
		/* the bottom of the V-cycle as one dense operator */
		int L = dense_level_of(D->prec, cs->N, D->nlev);
		D->dense_level = L;
		if (D->dense_level > 0)
			build_dense(D, L, cheb ? (gal ? 2 : 3) : 1, gal);
		if (cheb) {
			setup_cheb(D, cs, gal, gal ? mixed_deg0() : 2);
			D->cheb = 1;
		}
		if (gal) {
			for (int l = 1; l < D->nlev; l++)
				ggb_galop_free(&gal[l]);
			free(gal);
		}
		for (int l = 0; l < D->nlev; l++)
			host_level_free(&D->hl[l]);
		if (D->prec == GGB_PREC_MIXED) {
			for (int h = 0; h < 2; h++) {
				D->psi_prev[h] = xbuf(D, sizeof(double) * n, "fluid");
				ggbk_zero(D->psi_prev[h], sizeof(double) * n);
			}
		}
	}
	D->shadow_omega = falloc(n);
	D->shadow_band = falloc(n);
	D->shadow_psi = falloc(n);
	D->shadow_vel = falloc(3 * n);
	D->predict = 1;
	D->fix_its = -1;
	if (ggbk_check("fluid_init"))
		ggb_die("fluid_init: %s", ggbk_last_error());
	live_add(F);
	return 0;
}

void fluid_free(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	if (D) {
		live_remove(F);
		ggb_sync();
		for (int k = 0; k < D->nxbufs; k++)
			ggbk_free(D->xbufs[k]);
		free(D->xbufs);
		for (int l = 1; l < D->nlev; l++)
			cs_free(&D->coarse[l]);
		free(D->coarse);
		free(D->levels);
		free(D->hl);
		ggbk_host_free(D->hs);
		free(D->shadow_omega);
		free(D->shadow_band);
		free(D->shadow_psi);
		free(D->shadow_vel);
		free(D);
	}
	free(F->omega);
	free(F->omega_band);
	free(F->psi);
	free(F->vel);
	memset(F, 0, sizeof(*F));
}

/* ---------------- host mirrors ---------------- */

/* float host array -> device array of type T (through a staging buffer) */
static void put_f32(struct ggb_fluid_dev *D, void *dst, const float *src, size_t count)
{
	if (D->T == GGBK_F32) {
		upload(dst, src, 4 * count);
		return;
	}
	float *tmp = ggb_dmalloc(4 * count, "staging");
	upload(tmp, src, 4 * count);
	ggbk_convert(dst, GGBK_F64, tmp, GGBK_F32, count);
	ggb_sync();
	ggbk_free(tmp);
}

static void get_f32(struct ggb_fluid_dev *D, float *dst, const void *src, size_t count)
{
	if (D->T == GGBK_F32) {
		download(dst, src, 4 * count);
		return;
	}
	float *tmp = ggb_dmalloc(4 * count, "staging");
	ggbk_convert(tmp, GGBK_F32, src, GGBK_F64, count);
	download(dst, tmp, 4 * count);
	ggbk_free(tmp);
}

/* push caller-modified mirrors to the device (only arrays that changed) */
static void mirrors_pull(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t n = D->n;
	D->prev_valid = 0;
	if (memcmp(F->omega, D->shadow_omega, 4 * n)) {
		put_f32(D, D->omega, F->omega, n);
		memcpy(D->shadow_omega, F->omega, 4 * n);
	}
	if (memcmp(F->omega_band, D->shadow_band, 4 * n)) {
		put_f32(D, D->omega_band, F->omega_band, n);
		memcpy(D->shadow_band, F->omega_band, 4 * n);
	}
	if (memcmp(F->psi, D->shadow_psi, 4 * n)) {
		put_f32(D, D->psi, F->psi, n);
		memcpy(D->shadow_psi, F->psi, 4 * n);
	}
	if (memcmp(F->vel, D->shadow_vel, 12 * n)) {
		put_f32(D, D->vel, F->vel, 3 * n);
		memcpy(D->shadow_vel, F->vel, 12 * n);
		D->vmax_valid = 0;
	}
}

void fluid_sync_host(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t n = D->n;
	get_f32(D, F->omega, D->omega, n);
	get_f32(D, F->omega_band, D->omega_band, n);
	get_f32(D, F->psi, D->psi, n);
	get_f32(D, F->vel, D->vel, 3 * n);
	memcpy(D->shadow_omega, F->omega, 4 * n);
	memcpy(D->shadow_band, F->omega_band, 4 * n);
	memcpy(D->shadow_psi, F->psi, 4 * n);
	memcpy(D->shadow_vel, F->vel, 12 * n);
}

void fluid_sync_device(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t n = D->n;
	D->prev_valid = 0;
	put_f32(D, D->omega, F->omega, n);
	put_f32(D, D->omega_band, F->omega_band, n);
	put_f32(D, D->psi, F->psi, n);
	put_f32(D, D->vel, F->vel, 3 * n);
	D->vmax_valid = 0;
	memcpy(D->shadow_omega, F->omega, 4 * n);
	memcpy(D->shadow_band, F->omega_band, 4 * n);
	memcpy(D->shadow_psi, F->psi, 4 * n);
	memcpy(D->shadow_vel, F->vel, 12 * n);
}

static void *field_of(struct ggb_fluid_dev *D, int field, size_t *count)
{
	size_t n = D->n;
	switch (field) {
	case GGB_FIELD_OMEGA: *count = n; return D->omega;
	case GGB_FIELD_OMEGA_BAND: *count = n; return D->omega_band;
	case GGB_FIELD_PSI: *count = n; return D->psi;
	case GGB_FIELD_VEL: *count = 3 * n; return D->vel;
	case GGB_FIELD_DEP_B: *count = 3 * n; return D->dep_b;
	case GGB_FIELD_DEP_F: *count = 3 * n; return D->dep_f;
	default: *count = 0; return NULL;
	}
}

int fluid_get_f64(const Fluid *F, int field, double *out)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t cnt;
	void *src = field_of(D, field, &cnt);
	if (!src)
		return -1;
	if (D->T == GGBK_F64) {
		download(out, src, 8 * cnt);
		return 0;
	}
	double *tmp = ggb_dmalloc(8 * cnt, "staging");
	ggbk_convert(tmp, GGBK_F64, src, GGBK_F32, cnt);
	download(out, tmp, 8 * cnt);
	ggbk_free(tmp);
	return 0;
}

int fluid_set_f64(Fluid *F, int field, const double *in)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t cnt;
	void *dst = field_of(D, field, &cnt);
	if (!dst)
		return -1;
	if (field == GGB_FIELD_VEL)
		D->vmax_valid = 0;
	D->prev_valid = 0;
	if (D->T == GGBK_F64) {
		upload(dst, in, 8 * cnt);
		return 0;
	}
	double *tmp = ggb_dmalloc(8 * cnt, "staging");
	upload(tmp, in, 8 * cnt);
	ggbk_convert(dst, GGBK_F32, tmp, GGBK_F64, cnt);
	ggb_sync();
	ggbk_free(tmp);
	return 0;
}

/* ---------------- solver ---------------- */

static int mean_op(const struct ggb_fluid_dev *D) { return D->T == GGBK_F64 ? GGBK_FIN_MEAN_F64 : GGBK_FIN_MEAN_F32; }

/* z = M^-1 r and rz (fluid.c:148-160, 315-318, 344-348) */
static void precondition(struct ggb_fluid_dev *D, const void *r, int first)
{
	if (D->nlev > 0)
		ggbk_vcycle(D->levels, D->nlev, r, D->T, D->z, D->TM, D->s, D->red);
	else
		ggbk_jacobi_precond(D->g0, r, D->z, D->T, D->s, D->red);
	ggbk_fin(first ? GGBK_FIN_RZ_FIRST : GGBK_FIN_RZ, D->s, D->red, D->nb0, 0, 0, 0);
}

/* One PCG iteration after the first preconditioning (fluid.c:318-350): Ap and pAp, alpha,
 * then the update and the preconditioner.  With a multigrid hierarchy the update is fused
 * into the V-cycle's first pass and r alternates between two arrays. */
static void cg_iteration(struct ggb_fluid_dev *D, void *psi, int k)
{
	ggbk_grid g = D->g0;
	void *pin = D->pbuf[k & 1], *pout = D->pbuf[(k + 1) & 1];
	void *rin = D->rbuf[k & 1], *rout = D->rbuf[(k + 1) & 1];
	ggbk_cg_Ap(g, D->z, D->TZ, pin, pout, D->Ap, D->T, k == 0, D->s, D->red);
	ggbk_fin(GGBK_FIN_PAP, D->s, D->red, D->nb0, 0, 0, 0);
	ggbk_cg_fuse u = { psi, rin, pout, D->Ap };
	if (ggbk_vcycle_update(D->levels, D->nlev, &u, rout, D->T, D->z, D->TM, D->s, D->red) == 0) {
		ggbk_fin(GGBK_FIN_RZ, D->s, D->red, D->nb0, 0, 0, 0);
		return;
	}
	/* no hierarchy to fuse into: separate update in place */
	ggbk_cg_update(g, psi, rin, pout, D->Ap, D->T, D->s, D->red);
	ggbk_fin(GGBK_FIN_RR, D->s, D->red, D->nb0, 0, 0, 0);
	ggbk_d2d(rout, rin, D->tsz * (size_t) D->n);
	precondition(D, rout, 0);
}

static void read_scalars(struct ggb_fluid_dev *D)
{
	if (ggbk_d2h(D->hs, D->s, sizeof(ggbk_scalars)))
		ggb_die("device error: %s", ggbk_last_error());
	D->vmax_read = D->vmax_seq;
}

static void vmax_fold(struct ggb_fluid_dev *D)
{
	ggbk_fin(GGBK_FIN_VMAX, D->s, D->red, D->nb0, 0, 0, 0);
	D->vmax_seq++;
}

/* The passes after a solve (fluid.c:354 then :361-376): psi = 0 when b = 0, then
 * psi - mean(psi) into psi_alt fused into the velocity pass, and max |u|^2.  They read psi
 * and change no CG state, so a solve enqueues them before it reads its loop flag: the
 * read then waits for them too, and when the batch was short they run again after the
 * next one. */
static void post_solve_enqueue(struct ggb_fluid_dev *D)
{
	ggbk_mean_partial_bb0(D->g0, D->psi, D->T, D->s, D->red);
	ggbk_fin(mean_op(D), D->s, D->red, D->nb0, 0, 0, D->g0.total_area);
	ggbk_velocity(D->g0, D->psi, D->psi_alt, 1, D->vel, D->T, D->s, D->red);
	vmax_fold(D);
}

/* psi_alt (mean removed) becomes psi */
static void post_solve_finish(struct ggb_fluid_dev *D)
{
	void *t = D->psi;
	D->psi = D->psi_alt;
	D->psi_alt = t;
	D->vmax_valid = 1;
}

/* The mixed-precision PCG, the same iteration as the reference's (fluid.c:288-358) with the
 * work regrouped so that each iteration is one V-cycle: the V-cycle of r_k also yields
 * A z_k and the sums from which the fold after it (ggbk_cheb_fold) forms rz, beta, p.Ap =
 * z.Az + 2 beta z.Ap + beta^2 p.Ap (A symmetric), alpha and the residual norm after the
 * coming update, ||r - alpha Ap'||^2; the next V-cycle's first pass applies that update
 * (p' = z + beta p, Ap' = Az + beta Ap).  The stopping test thus sees each update's
 * residual before the update is applied, and the V-cycle after the last one is skipped.
 * Batches as in solve_on; a batch's last update is applied alone, and when the loop goes
 * on its residual is then preconditioned without an update. */
static void solve_cheb(struct ggb_fluid_dev *D, void *psi, int maxit, int post, double tol)
{
	ggbk_cg_cheb c;
	memset(&c, 0, sizeof(c));
	double *ap[2] = { D->Ap, D->pbuf[1] };
	c.psi = psi;
	c.p_in = c.p_out = D->pbuf[0];
	c.Az = D->Az;
	c.z = D->z;
	c.r = D->rbuf[0];
	/* the first V-cycle runs on cg_start's residual before anything is folded (cg_start
	 * opened it); its fold then starts the solve (bb, rr, the tolerance) and folds it */
	c.start = 1;
	c.nb0 = D->nb0;
	c.tol = tol;
	c.maxit = maxit;
	ggbk_vcycle_cheb(D->levels, D->nlev, &c, NULL, D->z, D->s, D->red);
	c.start = 0;
	int batch = D->predict > 1 ? D->predict : 1;
	int k = 0;
	for (;;) {
		for (int b = 0; b < batch && k < maxit; b++, k++) {
			c.update = 1;
			c.first = k == 0;
			c.r = D->rbuf[k & 1];
			c.r_out = D->rbuf[(k + 1) & 1];
			c.Ap_in = ap[k & 1];
			c.Ap_out = ap[(k + 1) & 1];
			if (b == batch - 1 || k == maxit - 1) {
				ggbk_cg_update_cheb(D->g0, &c, D->s);
				ggbk_fin(GGBK_FIN_CHEB_END, D->s, D->red, 0, 0, 0, 0);
			} else {
				ggbk_vcycle_cheb(D->levels, D->nlev, &c, NULL, D->z, D->s, D->red);
			}
		}
		if (post)
			post_solve_enqueue(D);
		read_scalars(D);
		if (!D->hs->active || k >= maxit)
			break;
		c.update = 0;
		c.r = D->rbuf[k & 1];
		c.Ap_in = ap[k & 1];
		ggbk_vcycle_cheb(D->levels, D->nlev, &c, NULL, D->z, D->s, D->red);
		batch = batch < 64 ? 2 * batch : 64;
	}
}

/* PCG for K psi = -area * omega on device arrays (reference fluid.c:288-358), psi
 * warm-started from its contents.  With sub_mean the pending mean of omega (in s->mean)
 * is subtracted by the first pass.  psi's mean is NOT removed. */
/* post: also the passes after the solve, enqueued before each read of the loop flag (psi
 * must be the state's psi; post_solve_finish completes them) */
static int solve_on(Fluid *F, void *omega, void *psi, int maxit, int sub_mean, int post)
{
	struct ggb_fluid_dev *D = F->dev;
	ggbk_grid g = D->g0;
	double tol = F->prm.cg_tol;
	if (D->fix_its >= 0) {
		/* a prescribed count: no residual is small enough to stop earlier */
		tol = 0.0;
		maxit = D->fix_its;
		D->predict = maxit > 0 ? maxit : 1;
	}
	ggbk_cg_start(g, omega, psi, D->rbuf[0], D->T, sub_mean, D->s, D->red, D->cheb);
	if (D->cheb) {
		solve_cheb(D, psi, maxit, post, tol);
		goto done;
	}
	ggbk_fin(GGBK_FIN_START, D->s, D->red, D->nb0, tol, maxit, 0);
	precondition(D, D->rbuf[0], 1);
	/* Iterations are enqueued in batches and the host reads the loop flag after each; the
	 * first batch is the previous solve's count, later batches grow geometrically.  Passes
	 * of iterations past convergence find the flag cleared and store nothing. */
	int batch = D->predict > 1 ? D->predict : 1;
	int k = 0;
	for (;;) {
		for (int b = 0; b < batch && k < maxit; b++, k++)
			cg_iteration(D, psi, k);
		if (post)
			post_solve_enqueue(D);
		read_scalars(D);
		if (!D->hs->active || k >= maxit)
			break;
		batch = batch < 64 ? 2 * batch : 64;
	}
done:
	if (!post)
		ggbk_cg_zero_if_bb0(g, psi, D->T, D->s);
	F->cg_iters = D->hs->it;
	F->cg_relres = D->hs->bb == 0.0 ? 0.0 : sqrt(D->hs->rr / D->hs->bb);
	D->predict = D->hs->it > 0 ? D->hs->it : 1;
	return D->hs->it;
}

static int solve_device(Fluid *F, int maxit, int sub_mean)
{
	return solve_on(F, F->dev->omega, F->dev->psi, maxit, sub_mean, 0);
}

/* psi -= mean(psi) fused into the velocity pass (fluid.c:354 then :361-376), after a
 * solve that did not enqueue them itself */
static void psi_mean_and_velocity(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	ggbk_mean_partial(D->g0, D->psi, D->T, D->red);
	ggbk_fin(mean_op(D), D->s, D->red, D->nb0, 0, 0, D->g0.total_area);
	ggbk_velocity(D->g0, D->psi, D->psi_alt, 1, D->vel, D->T, D->s, D->red);
	vmax_fold(D);
	post_solve_finish(D);
}

static void velocity_device(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	ggbk_velocity(D->g0, D->psi, NULL, 0, D->vel, D->T, D->s, D->red);
	vmax_fold(D);
	D->vmax_valid = 1;
}

int fluid_solve_psi_device(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	D->prev_valid = 0;
	int it = solve_device(F, F->prm.cg_maxit, 0);
	ggbk_remove_mean(D->g0, D->psi, D->T, D->s, D->red);
	if (ggbk_check("solve"))
		ggb_die("solve: %s", ggbk_last_error());
	return it;
}

int fluid_solve_psi(Fluid *F)
{
	mirrors_pull(F);
	int it = fluid_solve_psi_device(F);
	get_f32(F->dev, F->psi, F->dev->psi, F->dev->n);
	memcpy(F->dev->shadow_psi, F->psi, 4 * (size_t) F->dev->n);
	return it;
}

void fluid_compute_velocity(Fluid *F)
{
	mirrors_pull(F);
	velocity_device(F);
	get_f32(F->dev, F->vel, F->dev->vel, 3 * (size_t) F->dev->n);
	memcpy(F->dev->shadow_vel, F->vel, 12 * (size_t) F->dev->n);
}

void fluid_set_state(Fluid *F, const float *vel_initial, const float *vel_band)
{
	struct ggb_fluid_dev *D = F->dev;
	D->prev_valid = 0;
	ggbk_grid g = D->g0;
	size_t n = D->n;
	void *tmp = ggb_dmalloc(3 * D->tsz * n, "set_state");
	put_f32(D, tmp, vel_initial, 3 * n);
	ggbk_curl(g, tmp, D->omega, D->T);
	ggbk_remove_mean(g, D->omega, D->T, D->s, D->red);
	if (vel_band) {
		put_f32(D, tmp, vel_band, 3 * n);
		ggbk_curl(g, tmp, D->omega_band, D->T);
		ggbk_remove_mean(g, D->omega_band, D->T, D->s, D->red);
	} else {
		ggbk_zero(D->omega_band, D->tsz * n);
	}
	ggb_sync();
	ggbk_free(tmp);
	ggbk_zero(D->psi, D->tsz * n);
	int save = F->prm.cg_maxit;
	int cold = save * 20 > 20000 ? save * 20 : 20000;
	D->predict = 1;
	solve_device(F, cold, 0);
	D->predict = 1;
	psi_mean_and_velocity(F);
	fluid_sync_host(F);
	if (ggbk_check("fluid_set_state"))
		ggb_die("fluid_set_state: %s", ggbk_last_error());
}

/* reference fluid.c:424-478 */
void fluid_step(Fluid *F, double dt)
{
	struct ggb_fluid_dev *D = F->dev;
	ggbk_grid g = D->g0;
	int T = D->T;
	double twoOmega = 2.0 * F->prm.rotation;
	int bfecc = F->prm.bfecc;
	double relax = F->prm.forcing > 0 ? 1.0 - exp(-F->prm.forcing * dt) : 0.0;
	ggbk_departures(g, D->vel, D->omega, D->q0, D->dep_b, D->dep_f, T, dt, twoOmega, bfecc);
	ggbk_advect(g, D->q0, D->dep_b, D->q1, T);
	if (bfecc)
		ggbk_advect_bfecc(g, D->q1, D->q0, D->dep_f, D->q2, T);
	if (D->psi_prev[0] && !D->no_extrap) {
		/* psi(t + dt) by the Lagrange polynomial through the last three solutions, at times
		 * t, t - h1, t - h1 - h2 (fewer while the history fills up), formed in the
		 * limiter's pass (it touches none of psi's arrays) */
		double x = dt, h1 = D->dt_prev[0], h2 = D->dt_prev[1], c0 = 1.0, c1 = 0.0, c2 = 0.0;
		if (D->prev_valid >= 2) {
			c0 = (x + h1) * (x + h1 + h2) / (h1 * (h1 + h2));
			c1 = -x * (x + h1 + h2) / (h1 * h2);
			c2 = x * (x + h1) / ((h1 + h2) * h2);
		} else if (D->prev_valid == 1) {
			c0 = 1.0 + x / h1;
			c1 = -x / h1;
		}
		ggbk_limited_relax_extrap(g, D->q2, D->q0, D->dep_b, D->q1, bfecc, D->omega_band, D->omega, twoOmega,
					  relax, D->red, (double *) D->psi, D->psi_prev[0], D->psi_prev[1], c0, c1, c2);
		ggbk_fin(mean_op(D), D->s, D->red, D->nb0, 0, 0, g.total_area);
		double *t = D->psi_prev[0];
		D->psi_prev[0] = D->psi_prev[1];
		D->psi_prev[1] = t;
		D->dt_prev[1] = D->dt_prev[0];
		D->dt_prev[0] = dt;
		D->prev_valid = D->prev_valid < 2 ? D->prev_valid + 1 : 2;
	} else {
		ggbk_limited_relax(g, D->q2, D->q0, D->dep_b, D->q1, bfecc, D->omega_band, D->omega, T, twoOmega, relax,
				   D->red);
		ggbk_fin(mean_op(D), D->s, D->red, D->nb0, 0, 0, g.total_area);
	}
	solve_on(F, D->omega, D->psi, F->prm.cg_maxit, 1, 1);
	post_solve_finish(D);
	if (ggbk_check("fluid_step"))
		ggb_die("fluid_step: %s", ggbk_last_error());
	if (ggb_sync_mirrors())
		fluid_sync_host(F);
}

double fluid_max_speed(const Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	if (!D->vmax_valid) {
		ggbk_max_speed2(D->g0, D->vel, D->T, D->red);
		vmax_fold(D);
		D->vmax_valid = 1;
	}
	/* the last fluid_step's read of the loop flag came after its max-speed fold */
	if (D->vmax_read != D->vmax_seq)
		read_scalars(D);
	return sqrt(D->hs->vmax2);
}

double fluid_true_relres(Fluid *F)
{
	struct ggb_fluid_dev *D = F->dev;
	size_t n = D->n;
	double *kpsi = ggb_dmalloc(8 * n, "relres");
	ggbk_applyK(D->g0, D->psi, kpsi, D->T);
	double *hk = malloc(8 * n), *hw = malloc(8 * n);
	float *area = malloc(4 * n);
	if (!hk || !hw || !area)
		ggb_die("out of memory");
	download(hk, kpsi, 8 * n);
	ggbk_free(kpsi);
	fluid_get_f64(F, GGB_FIELD_OMEGA, hw);
	download(area, D->g0.area, 4 * n);
	double rr = 0, bb = 0;
	for (size_t i = 0; i < n; i++) {
		double b = -(double) area[i] * hw[i];
		double r = b - hk[i];
		rr += r * r;
		bb += b * b;
	}
	free(hk);
	free(hw);
	free(area);
	return bb > 0 ? sqrt(rr / bb) : 0.0;
}

int fluid_vcycle_f64(Fluid *F, const double *b, double *x)
{
	struct ggb_fluid_dev *D = F->dev;
	if (D->nlev < 1)
		return -1;
	size_t n = D->n;
	void *db = ggb_dmalloc(D->tsz * n, "vcycle");
	void *dz = ggb_dmalloc(D->tmsz * n, "vcycle");
	double *tmp = ggb_dmalloc(8 * n, "vcycle");
	upload(tmp, b, 8 * n);
	ggbk_convert(db, D->T, tmp, GGBK_F64, n);
	/* force the loop flag on for a stand-alone V-cycle */
	ggbk_scalars hs;
	memset(&hs, 0, sizeof(hs));
	hs.active = 1;
	hs.vc = 1;
	ggbk_scalars *ds = ggb_dmalloc(sizeof(ggbk_scalars), "vcycle");
	upload(ds, &hs, sizeof(hs));
	ggbk_red none = { NULL, 0, NULL };
	if (D->cheb)
		ggbk_vcycle_cheb(D->levels, D->nlev, NULL, db, dz, ds, none);
	else
		ggbk_vcycle(D->levels, D->nlev, db, D->T, dz, D->TM, ds, none);
	ggbk_convert(tmp, GGBK_F64, dz, D->TM, n);
	download(x, tmp, 8 * n);
	ggbk_free(db);
	ggbk_free(dz);
	ggbk_free(tmp);
	ggbk_free(ds);
	return ggbk_check("vcycle") ? -1 : 0;
}

/* ---------------- free operators (host arrays, reference semantics) ---------------- */

typedef void (*op_fn)(ggbk_grid, const void *, void *, int);

static void host_op(const CubeSphere *cs, op_fn fn, const void *in, size_t nin, void *out, size_t nout, size_t esz,
		    int T)
{
	ggbk_grid g = ggb_grid_dev(cs);
	void *din = ggb_dmalloc(esz * nin, "operator");
	void *dout = ggb_dmalloc(esz * nout, "operator");
	upload(din, in, esz * nin);
	fn(g, din, dout, T);
	download(out, dout, esz * nout);
	ggbk_free(din);
	ggbk_free(dout);
}

void fluid_curl(const CubeSphere *cs, const float *vel, float *omega)
{
	host_op(cs, ggbk_curl, vel, 3 * (size_t) cs->ncells, omega, cs->ncells, 4, GGBK_F32);
}

void fluid_divergence(const CubeSphere *cs, const float *vel, float *div)
{
	host_op(cs, ggbk_divergence, vel, 3 * (size_t) cs->ncells, div, cs->ncells, 4, GGBK_F32);
}

void fluid_laplacian(const CubeSphere *cs, const float *x, float *y)
{
	host_op(cs, ggbk_laplacian, x, cs->ncells, y, cs->ncells, 4, GGBK_F32);
}

int ggb_laplacian_f64(const CubeSphere *cs, const double *x, double *y)
{
	host_op(cs, ggbk_laplacian, x, cs->ncells, y, cs->ncells, 8, GGBK_F64);
	return 0;
}

int ggb_curl_f64(const CubeSphere *cs, const double *vel, double *out)
{
	host_op(cs, ggbk_curl, vel, 3 * (size_t) cs->ncells, out, cs->ncells, 8, GGBK_F64);
	return 0;
}

int ggb_divergence_f64(const CubeSphere *cs, const double *vel, double *out)
{
	host_op(cs, ggbk_divergence, vel, 3 * (size_t) cs->ncells, out, cs->ncells, 8, GGBK_F64);
	return 0;
}

int ggb_applyK_f64(const CubeSphere *cs, const double *x, double *y)
{
	ggbk_grid g = ggb_grid_dev(cs);
	size_t n = cs->ncells;
	void *din = ggb_dmalloc(8 * n, "operator");
	double *dout = ggb_dmalloc(8 * n, "operator");
	upload(din, x, 8 * n);
	ggbk_applyK(g, din, dout, GGBK_F64);
	download(y, dout, 8 * n);
	ggbk_free(din);
	ggbk_free(dout);
	return 0;
}

void fluid_remove_mean(const CubeSphere *cs, float *f)
{
	ggbk_grid g = ggb_grid_dev(cs);
	size_t n = cs->ncells;
	int nb = ggbk_nblocks((int) n);
	float *d = ggb_dmalloc(4 * n, "remove_mean");
	ggbk_scalars *s = ggb_dmalloc(sizeof(ggbk_scalars), "remove_mean");
	ggbk_red red = { NULL, nb, NULL };
	red.part = ggb_dmalloc(sizeof(double) * GGBK_NSLOTS * (size_t) nb, "remove_mean");
	ggbk_zero(s, sizeof(ggbk_scalars));
	upload(d, f, 4 * n);
	ggbk_remove_mean(g, d, GGBK_F32, s, red);
	download(f, d, 4 * n);
	ggbk_free(d);
	ggbk_free(s);
	ggbk_free(red.part);
}

/* fluid_remove_mean on device memory, allocating nothing (ggb_ext.h) */
int fluid_remove_mean_noalloc(const CubeSphere *cs, float *f_dev)
{
	if (cs->ncells > GGBK_NOALLOC_CELLS) {
		fprintf(stderr, "fluid_remove_mean_noalloc: %d cells exceed the scratch's %d\n", cs->ncells,
			GGBK_NOALLOC_CELLS);
		return -2;
	}
	if (!cs->dev) {
		fprintf(stderr, "fluid_remove_mean_noalloc: the grid's tables are not on the device\n");
		return -1;
	}
	int rc = ggbk_remove_mean_noalloc(cs->dev->g, f_dev);
	if (rc)
		fprintf(stderr, "fluid_remove_mean_noalloc: %s\n",
			rc == -2 ? "grid exceeds the scratch" : "device scratch unavailable or launch failed");
	return rc;
}

/* ---------------- discrete adjoint ----------------
 *
 * Gradient of a loss L(omega_T) with respect to the initial vorticity and the forcing
 * strength for a rollout of fluid_step with a prescribed dt schedule.  The forward pass
 * keeps (omega, psi) every K steps; the backward pass recomputes each K-step segment
 * from its checkpoint, keeping (omega_s, psi_s) of every step of the segment, then walks
 * the segment backwards through the adjoint kernels.  The Poisson solve is treated as
 * exact, so its adjoint is one solve with the same MG-PCG (K is symmetric).
 */
struct FluidAdjoint {
	Fluid *F;
	int K, nsteps, nck_cap;
	double *dts;           /* host [nsteps] */
	double *ckpt;          /* device [nck_cap][2][n]: omega, psi at steps 0, K, 2K, ... */
	double *tape;          /* device [K][2][n] */
	double *lam, *lq, *lq1, *lq2, *ldb, *ldf, *lvel, *lpsi, *om, *z, *dk;
	ggbk_adj_tr tr;        /* entries of the transposes (no floating-point atomics) */
	unsigned char *rslot;  /* device [maxnb][n]: slot of cell j in the list of its m-th neighbour */
	double sum_area;
	int *its;              /* host [nsteps + 1]: CG iterations of the last forward (set-up, steps) */
	int *fix;              /* host [nfix] or NULL: prescribed iteration counts */
	int nfix;
};

FluidAdjoint *fluid_adjoint_new(Fluid *F, int checkpoint_every)
{
	struct ggb_fluid_dev *D = F->dev;
	if (!D || D->T != GGBK_F64 || checkpoint_every < 1) {
		fprintf(stderr, "fluid_adjoint_new: needs an FP64-state fluid and K >= 1\n");
		return NULL;
	}
	const CubeSphere *cs = F->cs;
	size_t n = (size_t) D->n;
	int maxnb = D->g0.maxnb;
	unsigned char *rs = calloc((size_t) maxnb * n, 1);
	if (!rs)
		return NULL;
	for (size_t j = 0; j < n; j++)
		for (int m = 0; m < cs->nnb[j]; m++) {
			int i = cs->nbr[CS_MAXNB * j + m], k;
			for (k = 0; k < cs->nnb[i]; k++)
				if (cs->nbr[CS_MAXNB * (size_t) i + k] == (int) j)
					break;
			if (k == cs->nnb[i]) {
				fprintf(stderr, "fluid_adjoint_new: stencil not reciprocal at cells %zu, %d\n", j, i);
				free(rs);
				return NULL;
			}
			rs[(size_t) m * n + j] = (unsigned char) k;
		}
	FluidAdjoint *A = calloc(1, sizeof(*A));
	if (!A) {
		free(rs);
		return NULL;
	}
	A->F = F;
	A->K = checkpoint_every;
	for (size_t i = 0; i < n; i++)
		A->sum_area += cs->area[i];
	A->rslot = ggb_dmalloc((size_t) maxnb * n, "adjoint");
	upload(A->rslot, rs, (size_t) maxnb * n);
	free(rs);
	A->tape = ggb_dmalloc(sizeof(double) * 2 * n * (size_t) A->K, "adjoint tape");
	double **v1[] = { &A->lam, &A->lq, &A->lq1, &A->lq2, &A->lpsi, &A->om, &A->z };
	for (size_t k = 0; k < sizeof(v1) / sizeof(v1[0]); k++)
		*v1[k] = ggb_dmalloc(sizeof(double) * n, "adjoint");
	double **v3[] = { &A->ldb, &A->ldf, &A->lvel };
	for (size_t k = 0; k < sizeof(v3) / sizeof(v3[0]); k++)
		*v3[k] = ggb_dmalloc(sizeof(double) * 3 * n, "adjoint");
	A->dk = ggb_dmalloc(sizeof(double), "adjoint");
	/* 10 entries per cell (the departure points' taps), 2n targets (lq and lq2) */
	A->tr.cap_entries = 10 * (long long) n;
	A->tr.cap_targets = (int) (2 * n);
	A->tr.tgt = ggb_dmalloc(sizeof(int) * 10 * n, "adjoint");
	A->tr.val = ggb_dmalloc(sizeof(double) * 30 * n, "adjoint");
	A->tr.list = ggb_dmalloc(sizeof(int) * 10 * n, "adjoint");
	A->tr.cnt = ggb_dmalloc(sizeof(int) * 2 * n, "adjoint");
	A->tr.cursor = ggb_dmalloc(sizeof(int) * 2 * n, "adjoint");
	A->tr.start = ggb_dmalloc(sizeof(int) * (2 * n + 1), "adjoint");
	A->tr.bsum = ggb_dmalloc(sizeof(int) * (2 * n / 1024 + 2), "adjoint");
	return A;
}

void fluid_adjoint_free(FluidAdjoint *A)
{
	if (!A)
		return;
	ggb_sync();
	void *bufs[] = { A->ckpt, A->tape, A->lam, A->lq, A->lq1, A->lq2, A->ldb, A->ldf,
			 A->lvel, A->lpsi, A->om, A->z, A->dk, A->rslot, A->tr.tgt, A->tr.val,
			 A->tr.list, A->tr.cnt, A->tr.cursor, A->tr.start, A->tr.bsum };
	for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++)
		ggbk_free(bufs[k]);
	free(A->dts);
	free(A->its);
	free(A->fix);
	free(A);
}

static void adj_save(struct ggb_fluid_dev *D, double *dst)
{
	size_t n = (size_t) D->n;
	ggbk_d2d(dst, D->omega, 8 * n);
	ggbk_d2d(dst + n, D->psi, 8 * n);
}

int fluid_adjoint_forward(FluidAdjoint *A, const double *omega0, const double *dts, int nsteps, double *omega_T)
{
	Fluid *F = A->F;
	struct ggb_fluid_dev *D = F->dev;
	size_t n = (size_t) D->n;
	if (nsteps < 0)
		return -1;
	int nck = (nsteps + A->K - 1) / A->K;
	if (nck > A->nck_cap) {
		ggb_sync();
		ggbk_free(A->ckpt);
		A->ckpt = ggb_dmalloc(sizeof(double) * 2 * n * (size_t) nck, "adjoint checkpoints");
		A->nck_cap = nck;
	}
	free(A->dts);
	A->dts = malloc(sizeof(double) * (size_t) (nsteps ? nsteps : 1));
	if (!A->dts)
		return -1;
	memcpy(A->dts, dts, sizeof(double) * (size_t) nsteps);
	A->nsteps = nsteps;
	free(A->its);
	A->its = malloc(sizeof(int) * (size_t) (nsteps + 1));
	if (!A->its)
		return -1;
	if (A->fix && A->nfix != nsteps + 1) {
		fprintf(stderr, "fluid_adjoint_forward: %d prescribed iteration counts for %d solves\n", A->nfix,
			nsteps + 1);
		return -1;
	}
	/* set-up as fluid_set_state: omega - mean, cold solve, psi mean, velocity */
	upload(D->omega, omega0, 8 * n);
	ggbk_remove_mean(D->g0, D->omega, D->T, D->s, D->red);
	ggbk_zero(D->psi, 8 * n);
	int save = F->prm.cg_maxit;
	int cold = save * 20 > 20000 ? save * 20 : 20000;
	D->predict = 1;
	D->fix_its = A->fix ? A->fix[0] : -1;
	solve_device(F, cold, 0);
	A->its[0] = F->cg_iters;
	D->predict = 1;
	psi_mean_and_velocity(F);
	/* the solves start from the previous psi itself: a recomputed segment then repeats
	 * the forward pass bit for bit from its checkpoint */
	D->no_extrap = 1;
	for (int s = 0; s < nsteps; s++) {
		if (s % A->K == 0)
			adj_save(D, A->ckpt + 2 * n * (size_t) (s / A->K));
		D->fix_its = A->fix ? A->fix[s + 1] : -1;
		fluid_step(F, dts[s]);
		A->its[s + 1] = F->cg_iters;
	}
	D->fix_its = -1;
	D->no_extrap = 0;
	D->prev_valid = 0;
	if (omega_T)
		download(omega_T, D->omega, 8 * n);
	return ggbk_check("fluid_adjoint_forward") ? -1 : 0;
}

/* one backward step: lam = dL/d omega_{s+1} in, dL/d omega_s out; dL/dk accumulated */
static void adj_step(FluidAdjoint *A, double dt, const double *om_s, const double *psi_s)
{
	Fluid *F = A->F;
	struct ggb_fluid_dev *D = F->dev;
	ggbk_grid g = D->g0;
	size_t n = (size_t) D->n;
	int bfecc = F->prm.bfecc;
	double c = 2.0 * F->prm.rotation;
	double relax = F->prm.forcing > 0 ? 1.0 - exp(-F->prm.forcing * dt) : 0.0;
	double drelax = F->prm.forcing > 0 ? dt * exp(-F->prm.forcing * dt) : 0.0;
	/* forward intermediates of step s */
	ggbk_velocity(g, psi_s, NULL, 0, D->vel, D->T, D->s, D->red);
	ggbk_departures(g, D->vel, om_s, D->q0, D->dep_b, D->dep_f, D->T, dt, c, bfecc);
	ggbk_advect(g, D->q0, D->dep_b, D->q1, D->T);
	if (bfecc)
		ggbk_advect_bfecc(g, D->q1, D->q0, D->dep_f, D->q2, D->T);
	/* mean removal, relaxation, limiter */
	ggbk_adj_sum(g, A->lam, D->red);
	ggbk_fin(GGBK_FIN_MEAN_F64, D->s, D->red, D->nb0, 0, 0, g.total_area);
	ggbk_zero(A->lq, 8 * n);
	ggbk_zero(A->lq1, 8 * n);
	ggbk_zero(A->lq2, 8 * n);
	ggbk_zero(A->lvel, 24 * n);
	ggbk_adj_limiter(g, D->q0, D->q2, D->dep_b, D->omega_band, A->lam, D->s, relax, drelax, c, bfecc, &A->tr,
			 A->ldb, D->red);
	ggbk_fin(GGBK_FIN_MEAN_F64, D->s, D->red, D->nb0, 0, 0, g.total_area);
	ggbk_adj_accum(D->s, A->dk);
	if (ggbk_adj_transpose(&A->tr, 4 * (long long) n, 1, (int) (2 * n), (int) n, A->lq, A->lq2))
		ggb_die("fluid_adjoint_backward: transpose buffers too small");
	/* BFECC advections */
	if (bfecc) {
		ggbk_adj_bfecc_f(g, A->lq2, D->q1, D->dep_f, A->lq, &A->tr, A->ldf);
		if (ggbk_adj_transpose(&A->tr, 4 * (long long) n, 1, (int) n, (int) n, A->lq1, NULL))
			ggb_die("fluid_adjoint_backward: transpose buffers too small");
		ggbk_adj_advect_b(g, A->lq1, D->q0, D->dep_b, &A->tr, A->ldb);
		if (ggbk_adj_transpose(&A->tr, 4 * (long long) n, 1, (int) n, (int) n, A->lq, NULL))
			ggb_die("fluid_adjoint_backward: transpose buffers too small");
	}
	/* departure points, velocity, Poisson */
	ggbk_adj_departures(g, D->vel, A->ldb, A->ldf, dt, bfecc, &A->tr);
	if (ggbk_adj_transpose(&A->tr, 10 * (long long) n, 3, (int) n, (int) n, A->lvel, NULL))
		ggb_die("fluid_adjoint_backward: transpose buffers too small");
	ggbk_adj_velocity(g, A->lvel, A->rslot, A->lpsi, D->red);
	ggbk_fin(GGBK_FIN_MEAN_F64, D->s, D->red, D->nb0, 0, 0, g.total_area);
	ggbk_adj_rhs(g, A->lpsi, D->s, A->sum_area, A->om);
	solve_on(F, A->om, A->z, F->prm.cg_maxit, 0, 0);
	ggbk_adj_combine(g, A->lq, A->z, A->lam);
}

int fluid_adjoint_iterations(const FluidAdjoint *A, int *its)
{
	if (!A->its)
		return -1;
	if (its)
		memcpy(its, A->its, sizeof(int) * (size_t) (A->nsteps + 1));
	return A->nsteps + 1;
}

int fluid_adjoint_fix_iterations(FluidAdjoint *A, const int *its, int count)
{
	free(A->fix);
	A->fix = NULL;
	A->nfix = 0;
	if (!its)
		return 0;
	if (count < 1)
		return -1;
	A->fix = malloc(sizeof(int) * (size_t) count);
	if (!A->fix)
		return -1;
	for (int k = 0; k < count; k++)
		if (its[k] < 0) {
			free(A->fix);
			A->fix = NULL;
			return -1;
		}
	memcpy(A->fix, its, sizeof(int) * (size_t) count);
	A->nfix = count;
	return 0;
}

int fluid_adjoint_backward(FluidAdjoint *A, const double *g_omega_T, double *g_omega0, double *g_forcing)
{
	Fluid *F = A->F;
	struct ggb_fluid_dev *D = F->dev;
	ggbk_grid g = D->g0;
	size_t n = (size_t) D->n;
	if (!A->dts || !A->its)
		return -1;
	upload(A->lam, g_omega_T, 8 * n);
	ggbk_zero(A->dk, sizeof(double));
	ggbk_zero(A->z, 8 * n);
	int nseg = (A->nsteps + A->K - 1) / A->K;
	for (int seg = nseg - 1; seg >= 0; seg--) {
		int s0 = seg * A->K, s1 = s0 + A->K < A->nsteps ? s0 + A->K : A->nsteps;
		/* recompute the segment from its checkpoint, keeping every step's state */
		const double *ck = A->ckpt + 2 * n * (size_t) seg;
		ggbk_d2d(D->omega, ck, 8 * n);
		ggbk_d2d(D->psi, ck + n, 8 * n);
		velocity_device(F);
		D->no_extrap = 1;
		for (int s = s0; s < s1; s++) {
			adj_save(D, A->tape + 2 * n * (size_t) (s - s0));
			if (s + 1 < s1) {
				/* the forward's own iteration counts (from the same checkpoint the
				 * solves repeat bit for bit; with a prescribed schedule this keeps it) */
				D->fix_its = A->its[s + 1];
				fluid_step(F, A->dts[s]);
			}
		}
		D->fix_its = -1;
		D->no_extrap = 0;
		for (int s = s1 - 1; s >= s0; s--) {
			const double *t = A->tape + 2 * n * (size_t) (s - s0);
			adj_step(A, A->dts[s], t, t + n);
		}
	}
	/* leave the fluid in the rollout's initial state (checkpoint 0), with its velocity
	 * and maximum speed recomputed, rather than the mid-segment state of the last
	 * recomputation */
	ggbk_d2d(D->omega, A->ckpt, 8 * n);
	ggbk_d2d(D->psi, A->ckpt + n, 8 * n);
	velocity_device(F);
	D->prev_valid = 0;
	/* the control passes through set-up's mean removal */
	ggbk_adj_sum(g, A->lam, D->red);
	ggbk_fin(GGBK_FIN_MEAN_F64, D->s, D->red, D->nb0, 0, 0, g.total_area);
	ggbk_adj_project(g, A->lam, D->s, A->lq);
	if (g_omega0)
		download(g_omega0, A->lq, 8 * n);
	if (g_forcing)
		download(g_forcing, A->dk, sizeof(double));
	return ggbk_check("fluid_adjoint_backward") ? -1 : 0;
}
