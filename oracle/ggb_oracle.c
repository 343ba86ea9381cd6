/*
 * ggb_oracle.c -- CPU oracle for the per-step hot path.  TEST INFRASTRUCTURE ONLY:
 * it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
 * leg, as the checker.  The product library never links or calls it.
 *
 * A plain-C restatement of the reference algorithm (smcameron/gaseous-gusty-ball,
 * src/fluid.c, src/particles.c, src/render.c, and the sampling half of
 * src/cubesphere.c).  Each function cites the reference lines it follows and keeps
 * the reference's expression order and types, so that the float build
 * (ORC_REAL=float) reproduces the reference bit for bit when both are compiled with
 * -ffp-contract=off and run on one thread, and the double build (ORC_REAL=double)
 * is the FP64 oracle: the same algorithm with every state array stored in double.
 *
 * Geometry (cell centres, Voronoi stencil, gradient coefficients) is an INPUT: the
 * caller passes the per-level grid arrays (OGrid).  Stencil coefficients stay float
 * in both builds, exactly as the reference stores them (cubesphere.h:35-50).
 *
 * Reductions are sequential in cell order (the reference's OpenMP reductions with one
 * thread), so the oracle is deterministic.  Loops whose iterations are independent use
 * OpenMP.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   - float build vs the compiled reference (oracle/_ref): bitwise, every function.
 *   - double build: the reference's own invariants (spherical-harmonic Laplacian,
 *     PCG accuracy, curl(u) = omega, solid-body return), dense solves at N = 4/8,
 *     closed-form rotations, and agreement with the float build to float precision.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define MAXNB 8

/* ---------------- vector math (reference src/vec3.h) ---------------- */
typedef struct { double x, y, z; } v3d;

static inline v3d mk(double x, double y, double z) { v3d r = { x, y, z }; return r; }
static inline v3d vadd(v3d a, v3d b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3d vsub(v3d a, v3d b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3d vscl(v3d a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double vdot(v3d a, v3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3d vcross(v3d a, v3d b)
{
	return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen(v3d a) { return sqrt(vdot(a, a)); }
/* vec3.h:19-23: scale by the reciprocal length */
static inline v3d vnorm(v3d a)
{
	double l = vlen(a);
	return l > 0.0 ? vscl(a, 1.0 / l) : a;
}
/* vec3.h:25 */
static inline v3d vtan(v3d v, v3d n) { return vsub(v, vscl(n, vdot(v, n))); }
/* vec3.h:27 */
static inline double vangle(v3d a, v3d b) { return atan2(vlen(vcross(a, b)), vdot(a, b)); }
static inline v3d ldf(const float *f) { return mk(f[0], f[1], f[2]); }
static inline v3d ldr(const real *f) { return mk(f[0], f[1], f[2]); }
static inline v3d ldd(const double *f) { return mk(f[0], f[1], f[2]); }
static inline void str(real *f, v3d a) { f[0] = (real) a.x; f[1] = (real) a.y; f[2] = (real) a.z; }

/* vec3.h:33-42: great-circle move of p along tangent velocity v for time s */
static inline v3d gc_move(v3d p, v3d v, double s)
{
	v3d t = vtan(v, p);
	double sp = vlen(t);
	if (sp < 1e-300)
		return p;
	double th = sp * s;
	v3d r = vadd(vscl(p, cos(th)), vscl(t, sin(th) / sp));
	return vnorm(r);
}

/* ---------------- grid sampling (reference src/cubesphere.c:8-111) ---------------- */
static const v3d FC[6] = { {0, 0, 1}, {1, 0, 0}, {0, 0, -1}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0} };
static const v3d FR[6] = { {1, 0, 0}, {0, 0, -1}, {-1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0} };
static const v3d FU[6] = { {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1} };

typedef struct {
	int N, ncells;
	double h, total_area;
	const float *pos, *area;
	const unsigned char *nnb;
	const int *nbr;
	const float *w, *elen, *gc;
} OGrid;

/* cubesphere.c:12-20 */
static int face_of(v3d d)
{
	double ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
	if (ax >= ay && ax >= az)
		return d.x > 0 ? 1 : 3;
	if (ay >= az)
		return d.y > 0 ? 4 : 5;
	return d.z > 0 ? 0 : 2;
}

/* cubesphere.c:22-27 */
static v3d face_ab_dir(int f, double a, double b)
{
	return vnorm(vadd(FC[f], vadd(vscl(FR[f], tan(a)), vscl(FU[f], tan(b)))));
}

/* cubesphere.c:29-38 */
static void locate(const OGrid *g, v3d d, int *f, double *x, double *y)
{
	int face = face_of(d);
	double dc = vdot(d, FC[face]);
	double u = vdot(d, FR[face]) / dc;
	double v = vdot(d, FU[face]) / dc;
	*f = face;
	*x = (atan(u) + M_PI / 4) / g->h - 0.5;
	*y = (atan(v) + M_PI / 4) / g->h - 0.5;
}

static inline int cidx(const OGrid *g, int f, int i, int j) { return (f * g->N + j) * g->N + i; }

/* cubesphere.c:40-51 */
static int cell_of(const OGrid *g, v3d d)
{
	int f;
	double x, y;
	locate(g, d, &f, &x, &y);
	int i = (int) floor(x + 0.5), j = (int) floor(y + 0.5);
	if (i < 0) i = 0;
	if (i >= g->N) i = g->N - 1;
	if (j < 0) j = 0;
	if (j >= g->N) j = g->N - 1;
	return cidx(g, f, i, j);
}

/* cubesphere.c:53-66: ghost cells resolve to the real cell holding the extended centre */
static int tap(const OGrid *g, int f, int i, int j)
{
	if (i >= 0 && j >= 0 && i < g->N && j < g->N)
		return cidx(g, f, i, j);
	v3d d = face_ab_dir(f, -M_PI / 4 + (i + 0.5) * g->h, -M_PI / 4 + (j + 0.5) * g->h);
	return cell_of(g, d);
}

/* cubesphere.c:68-92 */
static void taps(const OGrid *g, v3d p, int idx[4], double wt[4])
{
	int f;
	double x, y;
	locate(g, p, &f, &x, &y);
	int i0 = (int) floor(x), j0 = (int) floor(y);
	double fx = x - i0, fy = y - j0;
	int N = g->N;
	if (i0 >= 0 && j0 >= 0 && i0 + 1 < N && j0 + 1 < N) {
		int b = cidx(g, f, i0, j0);
		idx[0] = b;
		idx[1] = b + 1;
		idx[2] = b + N;
		idx[3] = b + N + 1;
	} else {
		idx[0] = tap(g, f, i0, j0);
		idx[1] = tap(g, f, i0 + 1, j0);
		idx[2] = tap(g, f, i0, j0 + 1);
		idx[3] = tap(g, f, i0 + 1, j0 + 1);
	}
	wt[0] = (1 - fx) * (1 - fy);
	wt[1] = fx * (1 - fy);
	wt[2] = (1 - fx) * fy;
	wt[3] = fx * fy;
}

/* cubesphere.c:94-100 */
static double sample_s(const OGrid *g, const real *fld, v3d p)
{
	int idx[4];
	double wt[4];
	taps(g, p, idx, wt);
	return wt[0] * fld[idx[0]] + wt[1] * fld[idx[1]] + wt[2] * fld[idx[2]] + wt[3] * fld[idx[3]];
}

/* cubesphere.c:102-111 */
static v3d sample_v(const OGrid *g, const real *fld3, v3d p)
{
	int idx[4];
	double wt[4];
	taps(g, p, idx, wt);
	v3d r = mk(0, 0, 0);
	for (int k = 0; k < 4; k++)
		r = vadd(r, vscl(ldr(fld3 + 3 * idx[k]), wt[k]));
	return vtan(r, p);
}

static inline v3d gpos(const OGrid *g, int c) { return ldf(g->pos + 3 * c); }

/* Exported sampling helpers (for tap-table parity checks). */
void orc_taps(const OGrid *g, const double *p3, int *idx4, double *wt4)
{
	taps(g, mk(p3[0], p3[1], p3[2]), idx4, wt4);
}

int orc_cell_of(const OGrid *g, const double *p3) { return cell_of(g, mk(p3[0], p3[1], p3[2])); }

/* ---------------- fluid solver (reference src/fluid.c) ---------------- */
#define MG_SWEEPS 1
#define MG_OMEGA 0.8
#define MG_COARSE_SWEEPS 60
#define MG_CORR_SCALE 1.0

typedef struct {
	OGrid g;
	real *x, *b, *r;
	float *dinv;
	int *ptap;
	float *pw;
	int *rstart, *ridx;
	float *rw;
} OLevel;

typedef struct {
	OGrid g;
	double forcing, rotation;
	int bfecc;
	double cg_tol;
	int cg_maxit;
	int multigrid;
	int nlev;
	OLevel *lev;
	real *omega, *omega_band, *psi, *vel, *q0, *q1, *q2, *dep_b, *dep_f;
	real *r, *p, *Ap, *z;
	float *dinv;
	double *tmp;          /* per-cell terms for sequential reductions */
	int cg_iters;
	double cg_relres;
} OFluid;

static void *xcalloc(size_t n, size_t s)
{
	void *p = calloc(n ? n : 1, s);
	if (!p) {
		fprintf(stderr, "oracle: out of memory\n");
		exit(1);
	}
	return p;
}

/* sequential sum in cell order: what an OpenMP reduction computes on one thread */
static double seqsum(const double *t, int n)
{
	double s = 0;
	for (int i = 0; i < n; i++)
		s += t[i];
	return s;
}

/* fluid.c:277-286 */
static inline double applyK(const OGrid *g, const real *x, int i)
{
	const int *nb = g->nbr + MAXNB * i;
	const float *w = g->w + MAXNB * i;
	double xi = x[i], s = 0;
	for (int k = 0; k < g->nnb[i]; k++)
		s += w[k] * (xi - x[nb[k]]);
	return s;
}

static void dinv_of(const OGrid *g, float *dinv)
{
	/* fluid.c:58-63 and :182-187 */
	for (int i = 0; i < g->ncells; i++) {
		double s = 0;
		for (int k = 0; k < g->nnb[i]; k++)
			s += g->w[MAXNB * i + k];
		dinv[i] = (float) (1.0 / s);
	}
}

/* fluid.c:44-47: number of multigrid levels for face size N */
int orc_mg_nlev(int N)
{
	int nlev = 1;
	while ((N >> nlev) >= 4 && ((N >> (nlev - 1)) % 2) == 0)
		nlev++;
	return nlev;
}

/* fluid.c:41-98: per-level scratch and bilinear transfer operators (R = P^T as CSR). */
static void mg_setup(OFluid *F, const OGrid *levels)
{
	int nlev = orc_mg_nlev(F->g.N);
	F->nlev = nlev;
	F->lev = xcalloc(nlev, sizeof(OLevel));
	F->lev[0].g = F->g;
	F->lev[0].dinv = F->dinv;
	for (int l = 1; l < nlev; l++) {
		OLevel *L = &F->lev[l];
		L->g = levels[l];
		int n = L->g.ncells;
		L->x = xcalloc(n, sizeof(real));
		L->b = xcalloc(n, sizeof(real));
		L->r = xcalloc(n, sizeof(real));
		L->dinv = xcalloc(n, sizeof(float));
		dinv_of(&L->g, L->dinv);
		OLevel *Lf = &F->lev[l - 1];
		const OGrid *fine = &Lf->g;
		int nf = fine->ncells, nc = n;
		Lf->ptap = xcalloc(4 * (size_t) nf, sizeof(int));
		Lf->pw = xcalloc(4 * (size_t) nf, sizeof(float));
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < nf; i++) {
			int idx[4];
			double wt[4];
			taps(&L->g, gpos(fine, i), idx, wt);
			for (int k = 0; k < 4; k++) {
				Lf->ptap[4 * i + k] = idx[k];
				Lf->pw[4 * i + k] = (float) wt[k];
			}
		}
		Lf->rstart = xcalloc(nc + 1, sizeof(int));
		for (int i = 0; i < 4 * nf; i++)
			Lf->rstart[Lf->ptap[i] + 1]++;
		for (int c = 0; c < nc; c++)
			Lf->rstart[c + 1] += Lf->rstart[c];
		Lf->ridx = xcalloc(4 * (size_t) nf, sizeof(int));
		Lf->rw = xcalloc(4 * (size_t) nf, sizeof(float));
		int *cur = xcalloc(nc, sizeof(int));
		memcpy(cur, Lf->rstart, sizeof(int) * nc);
		for (int i = 0; i < nf; i++)
			for (int k = 0; k < 4; k++) {
				int c = Lf->ptap[4 * i + k];
				Lf->ridx[cur[c]] = i;
				Lf->rw[cur[c]] = Lf->pw[4 * i + k];
				cur[c]++;
			}
		free(cur);
	}
}

/* fluid.c:100-111 */
static void mg_smooth(const OGrid *g, const float *dinv, const real *b, real *x, real *r, int sweeps)
{
	int n = g->ncells;
	for (int s = 0; s < sweeps; s++) {
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < n; i++)
			r[i] = (real) (b[i] - applyK(g, x, i));
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < n; i++)
			x[i] += (real) (MG_OMEGA * dinv[i] * r[i]);
	}
}

/* fluid.c:113-146 */
static void mg_vcycle(OFluid *F, int l, const real *b, real *x, real *r)
{
	OLevel *L = &F->lev[l];
	const OGrid *g = &L->g;
	int n = g->ncells;
	memset(x, 0, sizeof(real) * n);
	if (l == F->nlev - 1) {
		mg_smooth(g, L->dinv, b, x, r, MG_COARSE_SWEEPS);
		return;
	}
	mg_smooth(g, L->dinv, b, x, r, MG_SWEEPS);
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++)
		r[i] = (real) (b[i] - applyK(g, x, i));
	OLevel *C = &F->lev[l + 1];
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < C->g.ncells; c++) {
		double acc = 0;
		for (int k = L->rstart[c]; k < L->rstart[c + 1]; k++)
			acc += L->rw[k] * r[L->ridx[k]];
		C->b[c] = (real) acc;
	}
	mg_vcycle(F, l + 1, C->b, C->x, C->r);
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		const int *t = L->ptap + 4 * i;
		const float *w = L->pw + 4 * i;
		x[i] += (real) (MG_CORR_SCALE * (w[0] * C->x[t[0]] + w[1] * C->x[t[1]] +
						  w[2] * C->x[t[2]] + w[3] * C->x[t[3]]));
	}
	mg_smooth(g, L->dinv, b, x, r, MG_SWEEPS);
}

/* fluid.c:149-160 */
static void precondition(OFluid *F, const real *r, real *z)
{
	int n = F->g.ncells;
	if (F->nlev > 0) {
		mg_vcycle(F, 0, r, z, F->Ap);
	} else {
		for (int i = 0; i < n; i++)
			z[i] = F->dinv[i] * r[i];
	}
}

/* Apply one V-cycle to b (exported so the GPU V-cycle can be checked on its own). */
void orc_vcycle(OFluid *F, const real *b, real *x)
{
	mg_vcycle(F, 0, b, x, F->Ap);
}

/* fluid.c:162-191 plus the coarse grids of mg_init.  levels[0] is the fine grid. */
OFluid *orc_fluid_new(const OGrid *levels, double forcing, double rotation, int bfecc,
		      double cg_tol, int cg_maxit, int multigrid)
{
	OFluid *F = xcalloc(1, sizeof(OFluid));
	F->g = levels[0];
	F->forcing = forcing;
	F->rotation = rotation;
	F->bfecc = bfecc;
	F->cg_tol = cg_tol;
	F->cg_maxit = cg_maxit;
	F->multigrid = multigrid;
	size_t n = F->g.ncells;
	F->omega = xcalloc(n, sizeof(real));
	F->omega_band = xcalloc(n, sizeof(real));
	F->psi = xcalloc(n, sizeof(real));
	F->vel = xcalloc(3 * n, sizeof(real));
	F->q0 = xcalloc(n, sizeof(real));
	F->q1 = xcalloc(n, sizeof(real));
	F->q2 = xcalloc(n, sizeof(real));
	F->dep_b = xcalloc(3 * n, sizeof(real));
	F->dep_f = xcalloc(3 * n, sizeof(real));
	F->r = xcalloc(n, sizeof(real));
	F->p = xcalloc(n, sizeof(real));
	F->Ap = xcalloc(n, sizeof(real));
	F->z = xcalloc(n, sizeof(real));
	F->dinv = xcalloc(n, sizeof(float));
	F->tmp = xcalloc(n, sizeof(double));
	dinv_of(&F->g, F->dinv);
	if (multigrid)
		mg_setup(F, levels);
	return F;
}

void orc_fluid_free(OFluid *F)
{
	if (!F)
		return;
	free(F->omega); free(F->omega_band); free(F->psi); free(F->vel);
	free(F->q0); free(F->q1); free(F->q2); free(F->dep_b); free(F->dep_f);
	free(F->r); free(F->p); free(F->Ap); free(F->z); free(F->dinv); free(F->tmp);
	for (int l = 0; l < F->nlev; l++) {
		OLevel *L = &F->lev[l];
		if (l > 0) {
			free(L->x); free(L->b); free(L->r); free(L->dinv);
		}
		free(L->ptap); free(L->pw); free(L->rstart); free(L->ridx); free(L->rw);
	}
	free(F->lev);
	free(F);
}

int orc_fluid_nlev(const OFluid *F) { return F->nlev; }
int orc_fluid_cg_iters(const OFluid *F) { return F->cg_iters; }
double orc_fluid_cg_relres(const OFluid *F) { return F->cg_relres; }

/* Transfer tables of level l (for tap-table parity checks of the product). */
void orc_fluid_transfer(const OFluid *F, int l, int *ptap, float *pw)
{
	const OLevel *L = &F->lev[l];
	memcpy(ptap, L->ptap, sizeof(int) * 4 * (size_t) L->g.ncells);
	memcpy(pw, L->pw, sizeof(float) * 4 * (size_t) L->g.ncells);
}

/* fluid.c:211-230 (operators take any-precision arrays via the storage type) */
void orc_divergence(const OGrid *g, const real *vel, real *div)
{
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d p = gpos(g, i);
		v3d ui = ldr(vel + 3 * i);
		double s = 0;
		for (int k = 0; k < g->nnb[i]; k++) {
			int j = g->nbr[MAXNB * i + k];
			v3d nrm = vnorm(vsub(gpos(g, j), p));
			v3d um = vadd(ui, ldr(vel + 3 * j));
			s += g->elen[MAXNB * i + k] * 0.5 * vdot(um, nrm);
		}
		div[i] = (real) (s / g->area[i]);
	}
}

/* fluid.c:232-250 */
void orc_curl(const OGrid *g, const real *vel, real *omega)
{
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d p = gpos(g, i);
		v3d wi = vcross(ldr(vel + 3 * i), p);
		double s = 0;
		for (int k = 0; k < g->nnb[i]; k++) {
			int j = g->nbr[MAXNB * i + k];
			v3d pj = gpos(g, j);
			v3d wj = vcross(ldr(vel + 3 * j), pj);
			v3d nrm = vnorm(vsub(pj, p));
			s += g->elen[MAXNB * i + k] * 0.5 * vdot(vadd(wi, wj), nrm);
		}
		omega[i] = (real) (s / g->area[i]);
	}
}

/* fluid.c:252-262 */
void orc_laplacian(const OGrid *g, const real *x, real *y)
{
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		double s = 0, xi = x[i];
		for (int k = 0; k < g->nnb[i]; k++)
			s += g->w[MAXNB * i + k] * (x[g->nbr[MAXNB * i + k]] - xi);
		y[i] = (real) (s / g->area[i]);
	}
}

/* Kx as a stored array (for the applyK parity checks). */
void orc_applyK(const OGrid *g, const real *x, double *y)
{
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < g->ncells; i++)
		y[i] = applyK(g, x, i);
}

/* fluid.c:264-275 */
void orc_remove_mean(const OGrid *g, real *f)
{
	int n = g->ncells;
	double s = 0;
	for (int i = 0; i < n; i++)
		s += (double) g->area[i] * f[i];
	real m = (real) (s / g->total_area);
	for (int i = 0; i < n; i++)
		f[i] -= m;
}

/* fluid.c:288-358 */
int orc_solve_psi(OFluid *F)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	real *psi = F->psi, *r = F->r, *p = F->p, *Ap = F->Ap, *z = F->z;
	double bb = 0, rr = 0, rz = 0;

	for (int i = 0; i < n; i++) {
		double b = -(double) g->area[i] * F->omega[i];
		double ri = b - applyK(g, psi, i);
		r[i] = (real) ri;
		bb += b * b;
		rr += ri * ri;
	}
	F->cg_iters = 0;
	if (bb == 0.0) {
		memset(psi, 0, sizeof(real) * n);
		F->cg_relres = 0;
		return 0;
	}
	double tol2 = F->cg_tol * F->cg_tol * bb;
	int it = 0;
	if (rr > tol2) {
		precondition(F, r, z);
		for (int i = 0; i < n; i++) {
			p[i] = z[i];
			rz += (double) r[i] * z[i];
		}
	}
	while (rr > tol2 && it < F->cg_maxit) {
		double pAp = 0;
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < n; i++) {
			double a = applyK(g, p, i);
			Ap[i] = (real) a;
			F->tmp[i] = a * p[i];
		}
		pAp = seqsum(F->tmp, n);
		if (pAp <= 0)
			break;
		double alpha = rz / pAp;
		rr = 0;
		for (int i = 0; i < n; i++) {
			psi[i] += (real) (alpha * p[i]);
			double ri = r[i] - alpha * Ap[i];
			r[i] = (real) ri;
			rr += ri * ri;
		}
		it++;
		if (rr <= tol2)
			break;
		precondition(F, r, z);
		double rznew = 0;
		for (int i = 0; i < n; i++)
			rznew += (double) r[i] * z[i];
		double beta = rznew / rz;
		rz = rznew;
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < n; i++)
			p[i] = (real) (z[i] + beta * p[i]);
	}
	orc_remove_mean(g, psi);
	F->cg_iters = it;
	F->cg_relres = sqrt(rr / bb);
	return it;
}

/* fluid.c:360-376 */
void orc_compute_velocity(OFluid *F)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d p = gpos(g, i);
		v3d gr = mk(0, 0, 0);
		double pi = F->psi[i];
		for (int k = 0; k < g->nnb[i]; k++) {
			int j = g->nbr[MAXNB * i + k];
			gr = vadd(gr, vscl(ldf(g->gc + 3 * (MAXNB * i + k)), F->psi[j] - pi));
		}
		str(F->vel + 3 * i, vcross(p, gr));
	}
}

/* fluid.c:378-395; the velocity inputs are float arrays, as in the reference */
void orc_fluid_set_state(OFluid *F, const float *vel_initial, const float *vel_band)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	real *tmp = xcalloc(3 * (size_t) n, sizeof(real));
	for (int i = 0; i < 3 * n; i++)
		tmp[i] = vel_initial[i];
	orc_curl(g, tmp, F->omega);
	orc_remove_mean(g, F->omega);
	if (vel_band) {
		for (int i = 0; i < 3 * n; i++)
			tmp[i] = vel_band[i];
		orc_curl(g, tmp, F->omega_band);
		orc_remove_mean(g, F->omega_band);
	} else {
		memset(F->omega_band, 0, sizeof(real) * n);
	}
	free(tmp);
	memset(F->psi, 0, sizeof(real) * n);
	int save = F->cg_maxit;
	F->cg_maxit = save * 20 > 20000 ? save * 20 : 20000;
	orc_solve_psi(F);
	F->cg_maxit = save;
	orc_compute_velocity(F);
}

/* fluid.c:397-414 */
static void departures(OFluid *F, double dt, int forward_too)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d p = gpos(g, i);
		v3d u = ldr(F->vel + 3 * i);
		v3d pm = gc_move(p, u, -0.5 * dt);
		v3d um = vtan(sample_v(g, F->vel, pm), p);
		str(F->dep_b + 3 * i, gc_move(p, um, -dt));
		if (forward_too) {
			pm = gc_move(p, u, 0.5 * dt);
			um = vtan(sample_v(g, F->vel, pm), p);
			str(F->dep_f + 3 * i, gc_move(p, um, dt));
		}
	}
}

/* fluid.c:416-422 */
static void advect(const OGrid *g, const real *src, const real *dep, real *dst)
{
	int n = g->ncells;
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++)
		dst[i] = (real) sample_s(g, src, ldr(dep + 3 * i));
}

/* fluid.c:424-478 */
void orc_fluid_step(OFluid *F, double dt)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	double twoOmega = 2.0 * F->rotation;
	int bfecc = F->bfecc;

	departures(F, dt, bfecc);

	real *q = F->q0;
	for (int i = 0; i < n; i++)
		q[i] = (real) (F->omega[i] + twoOmega * g->pos[3 * i + 1]);

	real *qn = F->q1;
	if (!bfecc) {
		advect(g, q, F->dep_b, qn);
	} else {
		real *q1 = F->q1, *q2 = F->q2;
		advect(g, q, F->dep_b, q1);
		advect(g, q1, F->dep_f, q2);
		for (int i = 0; i < n; i++)
			q2[i] = q[i] + (real) 0.5f * (q[i] - q2[i]);
		#pragma omp parallel for schedule(static)
		for (int i = 0; i < n; i++) {
			int idx[4];
			double wt[4];
			taps(g, ldr(F->dep_b + 3 * i), idx, wt);
			double v = 0, lo = 1e30, hi = -1e30;
			for (int k = 0; k < 4; k++) {
				v += wt[k] * q2[idx[k]];
				double o = q[idx[k]];
				if (o < lo) lo = o;
				if (o > hi) hi = o;
			}
			if (v < lo) v = lo;
			if (v > hi) v = hi;
			q1[i] = (real) v;
		}
		qn = q1;
	}

	double relax = F->forcing > 0 ? 1.0 - exp(-F->forcing * dt) : 0.0;
	for (int i = 0; i < n; i++) {
		double w = qn[i] - twoOmega * g->pos[3 * i + 1];
		w += relax * (F->omega_band[i] - w);
		F->omega[i] = (real) w;
	}
	orc_remove_mean(g, F->omega);
	orc_solve_psi(F);
	orc_compute_velocity(F);
}

/* fluid.c:480-491 */
double orc_fluid_max_speed(const OFluid *F)
{
	int n = F->g.ncells;
	double m = 0;
	for (int i = 0; i < n; i++) {
		const real *v = F->vel + 3 * i;
		double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
		if (s > m) m = s;
	}
	return sqrt(m);
}

/* State access: which = 0 omega, 1 omega_band, 2 psi, 3 vel (3n), 4 dep_b (3n),
 * 5 dep_f (3n).  Values cross the boundary as double. */
static real *field_ptr(OFluid *F, int which, size_t *len)
{
	size_t n = F->g.ncells;
	switch (which) {
	case 0: *len = n; return F->omega;
	case 1: *len = n; return F->omega_band;
	case 2: *len = n; return F->psi;
	case 3: *len = 3 * n; return F->vel;
	case 4: *len = 3 * n; return F->dep_b;
	case 5: *len = 3 * n; return F->dep_f;
	default: *len = 0; return NULL;
	}
}

int orc_fluid_get(OFluid *F, int which, double *out)
{
	size_t len;
	real *p = field_ptr(F, which, &len);
	if (!p)
		return -1;
	for (size_t i = 0; i < len; i++)
		out[i] = p[i];
	return 0;
}

int orc_fluid_set(OFluid *F, int which, const double *in)
{
	size_t len;
	real *p = field_ptr(F, which, &len);
	if (!p)
		return -1;
	for (size_t i = 0; i < len; i++)
		p[i] = (real) in[i];
	return 0;
}

/* ---------------- discrete adjoint of the fluid step ----------------
 *
 * Reverse-mode derivative of n steps of orc_fluid_step (this build's FP64 state) with a
 * prescribed dt schedule, for the loss gradient with respect to the initial vorticity
 * and to the forcing strength k (relax = 1 - exp(-k dt), fluid.c:466).  The forward
 * rollout keeps (omega_n, psi_n) for every step (a full tape); the backward sweep
 * recomputes each step's intermediates from them.  The Poisson solve is treated as exact
 * (psi = Pi_A K^+ (-A omega) with Pi_A the area-weighted mean removal), so its adjoint
 * is one more solve with the same PCG; the limiter's active set (fluid.c:455-462) is
 * re-evaluated from the recomputed forward values, ties resolved to the first tap as the
 * forward loop does.  The control omega0 enters through the mean removal of set-up.
 */

/* x(p), y(p) of locate() and their gradients with respect to the 3-vector p on face f */
static void locate_grad(const OGrid *g, v3d p, int f, v3d *gx, v3d *gy)
{
	double dc = vdot(p, FC[f]);
	double u = vdot(p, FR[f]) / dc, v = vdot(p, FU[f]) / dc;
	v3d du = vscl(vsub(FR[f], vscl(FC[f], u)), 1.0 / dc);
	v3d dv = vscl(vsub(FU[f], vscl(FC[f], v)), 1.0 / dc);
	*gx = vscl(du, 1.0 / (g->h * (1 + u * u)));
	*gy = vscl(dv, 1.0 / (g->h * (1 + v * v)));
}

/* taps() plus the gradients of the four bilinear weights with respect to p */
static void taps_grad(const OGrid *g, v3d p, int idx[4], double wt[4], v3d dw[4])
{
	taps(g, p, idx, wt);
	int f;
	double x, y;
	locate(g, p, &f, &x, &y);
	double fx = x - floor(x), fy = y - floor(y);
	v3d gx, gy;
	locate_grad(g, p, f, &gx, &gy);
	dw[0] = vadd(vscl(gx, -(1 - fy)), vscl(gy, -(1 - fx)));
	dw[1] = vadd(vscl(gx, 1 - fy), vscl(gy, -fx));
	dw[2] = vadd(vscl(gx, -fy), vscl(gy, 1 - fx));
	dw[3] = vadd(vscl(gx, fy), vscl(gy, fx));
}

/* adjoint of out = gc_move(p, v, s) with respect to v (p fixed): returns lambda_v */
static v3d gc_move_adj(v3d p, v3d v, double s, v3d lo)
{
	v3d t = vtan(v, p);
	double sp = vlen(t);
	if (sp < 1e-300)
		return mk(0, 0, 0);
	double th = sp * s, sn = sin(th), cs = cos(th);
	v3d r = vadd(vscl(p, cs), vscl(t, sn / sp));
	double l = vlen(r);
	v3d out = vscl(r, 1.0 / l);
	v3d lr = vscl(vsub(lo, vscl(out, vdot(out, lo))), 1.0 / l);
	double tl = vdot(t, lr);
	v3d lt = vscl(lr, sn / sp);
	double lth = -sn * vdot(p, lr) + cs / sp * tl;
	double lsp = -sn / (sp * sp) * tl + s * lth;
	lt = vadd(lt, vscl(t, lsp / sp));
	return vtan(lt, p);
}

/* adjoint of one departure point (fluid.c:404-412) with steps s1, s2: adds to lvel */
static void departure_adj(const OGrid *g, const real *vel, int i, double s1, double s2, v3d ldep, double *lvel)
{
	v3d p = gpos(g, i), u = ldr(vel + 3 * i);
	v3d pm = gc_move(p, u, s1);
	int idx[4];
	double wt[4];
	v3d dw[4];
	taps_grad(g, pm, idx, wt, dw);
	v3d r = mk(0, 0, 0);
	for (int k = 0; k < 4; k++)
		r = vadd(r, vscl(ldr(vel + 3 * idx[k]), wt[k]));
	v3d sv = vtan(r, pm);
	v3d um = vtan(sv, p);
	v3d lum = gc_move_adj(p, um, s2, ldep);
	v3d lsv = vtan(lum, p);
	v3d lr = vtan(lsv, pm);
	v3d lpm = vsub(vscl(lsv, -vdot(pm, r)), vscl(r, vdot(pm, lsv)));
	for (int k = 0; k < 4; k++) {
		double *o = lvel + 3 * idx[k];
		o[0] += wt[k] * lr.x;
		o[1] += wt[k] * lr.y;
		o[2] += wt[k] * lr.z;
		lpm = vadd(lpm, vscl(dw[k], vdot(ldr(vel + 3 * idx[k]), lr)));
	}
	v3d lu = gc_move_adj(p, u, s1, lpm);
	lvel[3 * i] += lu.x;
	lvel[3 * i + 1] += lu.y;
	lvel[3 * i + 2] += lu.z;
}

/* lsrc[taps] += w * ldst (transpose of advect) and ldep += sum dw * src[taps] * ldst */
static void advect_adj(const OGrid *g, const real *src, const real *dep, const double *ldst, double *lsrc,
		       double *ldep)
{
	int n = g->ncells;
	for (int i = 0; i < n; i++) {
		int idx[4];
		double wt[4];
		v3d dw[4];
		taps_grad(g, ldr(dep + 3 * i), idx, wt, dw);
		v3d d = mk(0, 0, 0);
		for (int k = 0; k < 4; k++) {
			lsrc[idx[k]] += wt[k] * ldst[i];
			d = vadd(d, vscl(dw[k], (double) src[idx[k]] * ldst[i]));
		}
		ldep[3 * i] += d.x;
		ldep[3 * i + 1] += d.y;
		ldep[3 * i + 2] += d.z;
	}
}

typedef struct {
	int nsteps;
	double *dts;
	real *om, *ps;     /* [nsteps][n] omega_n, psi_n */
} OTape;

/* the tape of the last orc_adj_forward of each fluid (one fluid at a time is enough
 * for the checker; kept beside the fluid by pointer) */
typedef struct { OFluid *F; OTape T; } OTapeSlot;
static OTapeSlot g_tapes[8];

static OTape *tape_of(OFluid *F, int create)
{
	for (int k = 0; k < 8; k++)
		if (g_tapes[k].F == F)
			return &g_tapes[k].T;
	if (!create)
		return NULL;
	for (int k = 0; k < 8; k++)
		if (!g_tapes[k].F) {
			g_tapes[k].F = F;
			memset(&g_tapes[k].T, 0, sizeof(OTape));
			return &g_tapes[k].T;
		}
	return NULL;
}

void orc_adj_release(OFluid *F)
{
	for (int k = 0; k < 8; k++)
		if (g_tapes[k].F == F) {
			free(g_tapes[k].T.dts);
			free(g_tapes[k].T.om);
			free(g_tapes[k].T.ps);
			memset(&g_tapes[k], 0, sizeof(g_tapes[k]));
		}
}

void orc_fluid_set_forcing(OFluid *F, double forcing) { F->forcing = forcing; }

/* Forward rollout from the control omega0: omega = omega0 - mean, psi solved from zero,
 * velocity; then nsteps steps with the given dts.  Records the tape. */
int orc_adj_forward(OFluid *F, const double *omega0, const double *dts, int nsteps)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	orc_adj_release(F);
	OTape *T = tape_of(F, 1);
	if (!T || nsteps < 0)
		return -1;
	T->nsteps = nsteps;
	T->dts = xcalloc(nsteps, sizeof(double));
	T->om = xcalloc((size_t) nsteps * n, sizeof(real));
	T->ps = xcalloc((size_t) nsteps * n, sizeof(real));
	memcpy(T->dts, dts, sizeof(double) * nsteps);
	for (int i = 0; i < n; i++)
		F->omega[i] = (real) omega0[i];
	orc_remove_mean(g, F->omega);
	memset(F->psi, 0, sizeof(real) * n);
	orc_solve_psi(F);
	orc_compute_velocity(F);
	for (int s = 0; s < nsteps; s++) {
		memcpy(T->om + (size_t) s * n, F->omega, sizeof(real) * n);
		memcpy(T->ps + (size_t) s * n, F->psi, sizeof(real) * n);
		orc_fluid_step(F, dts[s]);
	}
	return 0;
}

/* Backward sweep: gT = dL/d omega_T in; g_omega0 = dL/d omega0 and *g_forcing = dL/dk
 * out.  Leaves the fluid's state arrays overwritten by the recomputation. */
int orc_adj_backward(OFluid *F, const double *gT, double *g_omega0, double *g_forcing)
{
	const OGrid *g = &F->g;
	int n = g->ncells;
	OTape *T = tape_of(F, 0);
	if (!T)
		return -1;
	double c = 2.0 * F->rotation;
	double *lam = xcalloc(n, sizeof(double)), *lpp = xcalloc(n, sizeof(double));
	double *lq = xcalloc(n, sizeof(double)), *lq1 = xcalloc(n, sizeof(double));
	double *lq2 = xcalloc(n, sizeof(double)), *lq2p = xcalloc(n, sizeof(double));
	double *ldbw = xcalloc(3 * (size_t) n, sizeof(double)), *ldfw = xcalloc(3 * (size_t) n, sizeof(double));
	double *lvel = xcalloc(3 * (size_t) n, sizeof(double)), *lpsi = xcalloc(n, sizeof(double));
	real *z = xcalloc(n, sizeof(real)), *rhs_om = xcalloc(n, sizeof(real));
	memcpy(lam, gT, sizeof(double) * n);
	double dk = 0;
	for (int s = T->nsteps - 1; s >= 0; s--) {
		double dt = T->dts[s];
		/* recompute step s's forward intermediates */
		memcpy(F->omega, T->om + (size_t) s * n, sizeof(real) * n);
		memcpy(F->psi, T->ps + (size_t) s * n, sizeof(real) * n);
		orc_compute_velocity(F);
		departures(F, dt, F->bfecc);
		real *q = F->q0, *q1 = F->q1, *q2 = F->q2;
		for (int i = 0; i < n; i++)
			q[i] = (real) (F->omega[i] + c * g->pos[3 * i + 1]);
		advect(g, q, F->dep_b, q1);
		if (F->bfecc) {
			advect(g, q1, F->dep_f, q2);
			for (int i = 0; i < n; i++)
				q2[i] = q[i] + (real) 0.5f * (q[i] - q2[i]);
		}
		double relax = F->forcing > 0 ? 1.0 - exp(-F->forcing * dt) : 0.0;
		double drelax = F->forcing > 0 ? dt * exp(-F->forcing * dt) : 0.0;
		/* mean removal (fluid.c:475): lpp = Pi_A^T lam */
		double sum = 0;
		for (int i = 0; i < n; i++)
			sum += lam[i];
		double m = sum / g->total_area;
		for (int i = 0; i < n; i++)
			lpp[i] = lam[i] - (double) g->area[i] * m;
		memset(lq, 0, sizeof(double) * n);
		memset(lq1, 0, sizeof(double) * n);
		memset(lq2, 0, sizeof(double) * n);
		memset(ldbw, 0, sizeof(double) * 3 * n);
		memset(ldfw, 0, sizeof(double) * 3 * n);
		memset(lvel, 0, sizeof(double) * 3 * n);
		/* relaxation and limiter (fluid.c:449-474) */
		for (int i = 0; i < n; i++) {
			int idx[4];
			double wt[4];
			v3d dw[4];
			taps_grad(g, ldr(F->dep_b + 3 * i), idx, wt, dw);
			const real *src = F->bfecc ? q2 : q;
			double v = 0, lo = 1e30, hi = -1e30;
			int klo = 0, khi = 0;
			for (int k = 0; k < 4; k++) {
				v += wt[k] * src[idx[k]];
				double o = q[idx[k]];
				if (o < lo) { lo = o; klo = k; }
				if (o > hi) { hi = o; khi = k; }
			}
			int state = 0;
			double qn = v;
			if (F->bfecc) {
				if (v < lo) { qn = lo; state = 1; }
				if (qn > hi) { qn = hi; state = 2; }
			}
			double w = qn - c * g->pos[3 * i + 1];
			dk += lpp[i] * ((double) F->omega_band[i] - w) * drelax;
			double lqn = (1 - relax) * lpp[i];
			if (state == 1) {
				lq[idx[klo]] += lqn;
			} else if (state == 2) {
				lq[idx[khi]] += lqn;
			} else {
				double *lt = F->bfecc ? lq2 : lq;
				v3d d = mk(0, 0, 0);
				for (int k = 0; k < 4; k++) {
					lt[idx[k]] += wt[k] * lqn;
					d = vadd(d, vscl(dw[k], (double) src[idx[k]] * lqn));
				}
				ldbw[3 * i] += d.x;
				ldbw[3 * i + 1] += d.y;
				ldbw[3 * i + 2] += d.z;
			}
		}
		if (F->bfecc) {
			/* q2 = q + 0.5 (q - A_f q1) (fluid.c:446-448), then q1 = A_b q */
			for (int i = 0; i < n; i++) {
				lq[i] += 1.5 * lq2[i];
				lq2p[i] = -0.5 * lq2[i];
			}
			advect_adj(g, q1, F->dep_f, lq2p, lq1, ldfw);
			advect_adj(g, q, F->dep_b, lq1, lq, ldbw);
		}
		/* departures (fluid.c:397-414) */
		for (int i = 0; i < n; i++) {
			departure_adj(g, F->vel, i, -0.5 * dt, -dt, ldd(ldbw + 3 * i), lvel);
			if (F->bfecc)
				departure_adj(g, F->vel, i, 0.5 * dt, dt, ldd(ldfw + 3 * i), lvel);
		}
		/* velocity (fluid.c:360-376): vel_i = p_i x sum_k gc_ik (psi_j - psi_i) */
		memset(lpsi, 0, sizeof(double) * n);
		for (int i = 0; i < n; i++) {
			v3d mvec = vcross(ldd(lvel + 3 * i), gpos(g, i));
			for (int k = 0; k < g->nnb[i]; k++) {
				int j = g->nbr[MAXNB * i + k];
				double cik = vdot(ldf(g->gc + 3 * (MAXNB * i + k)), mvec);
				lpsi[j] += cik;
				lpsi[i] -= cik;
			}
		}
		/* Poisson: psi = Pi_A K^+ (-A omega): lam_omega = -A K^+ Pi_A^T lpsi.  K^+ ignores
		 * the constant component of its argument; it is removed explicitly (Pi_A uses the
		 * geometric total area, which differs from the sum of the float areas in the last
		 * bits), so the solve sees a consistent right-hand side. */
		double ps = 0, sa = 0;
		for (int i = 0; i < n; i++) {
			ps += lpsi[i];
			sa += g->area[i];
		}
		double pm = ps / g->total_area, cm = ps * (1.0 - sa / g->total_area) / n;
		for (int i = 0; i < n; i++)
			rhs_om[i] = (real) (-(lpsi[i] - (double) g->area[i] * pm - cm) / (double) g->area[i]);
		real *so = F->omega, *sp = F->psi;
		F->omega = rhs_om;
		F->psi = z;
		orc_solve_psi(F);
		F->omega = so;
		F->psi = sp;
		for (int i = 0; i < n; i++)
			lam[i] = lq[i] - (double) g->area[i] * z[i];
	}
	/* set-up mean removal of the control */
	double sum = 0;
	for (int i = 0; i < n; i++)
		sum += lam[i];
	double m = sum / g->total_area;
	for (int i = 0; i < n; i++)
		g_omega0[i] = lam[i] - (double) g->area[i] * m;
	*g_forcing = dk;
	free(lam); free(lpp); free(lq); free(lq1); free(lq2); free(lq2p);
	free(ldbw); free(ldfw); free(lvel); free(lpsi); free(z); free(rhs_om);
	return 0;
}

/* ---------------- particles (reference src/particles.c) ---------------- */

/* particles.c:8-31 (positions only; colours start at 1) */
void orc_particles_init(int n, real *pos, real *prev, float *col)
{
	const double golden = M_PI * (3.0 - sqrt(5.0));
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		double y = 1.0 - 2.0 * (i + 0.5) / n;
		double r = sqrt(1.0 - y * y);
		double phi = golden * i;
		v3d p = mk(r * cos(phi), y, r * sin(phi));
		str(pos + 3 * i, p);
		str(prev + 3 * i, p);
	}
	for (int i = 0; i < 4 * n; i++)
		col[i] = 1.0f;
}

/* particles.c:41-61 (vec3.h:46-53 for lon/lat) */
void orc_particles_color(int n, const real *pos, float *col, const float *img, int w, int h)
{
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d d = ldr(pos + 3 * i);
		double lon = atan2(d.x, d.z);
		double yy = d.y;
		if (yy > 1.0) yy = 1.0;
		if (yy < -1.0) yy = -1.0;
		double lat = asin(yy);
		double x = (lon + M_PI) / (2 * M_PI) * w - 0.5;
		double y = (M_PI / 2 - lat) / M_PI * h - 0.5;
		int x0 = (int) floor(x), y0 = (int) floor(y);
		double fx = x - x0, fy = y - y0;
		int xs[2] = { ((x0 % w) + w) % w, (((x0 + 1) % w) + w) % w };
		int ys[2] = { y0 < 0 ? 0 : (y0 >= h ? h - 1 : y0),
			      y0 + 1 < 0 ? 0 : (y0 + 1 >= h ? h - 1 : y0 + 1) };
		double wts[4] = { (1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy };
		for (int c = 0; c < 4; c++) {
			double v = 0;
			for (int k = 0; k < 4; k++)
				v += wts[k] * img[((size_t) ys[k >> 1] * w + xs[k & 1]) * 4 + c];
			col[4 * i + c] = (float) v;
		}
	}
}

/* particles.c:63-74 */
void orc_particles_advect(const OGrid *g, int n, real *pos, real *prev, const real *vel, double dt)
{
	#pragma omp parallel for schedule(static)
	for (int i = 0; i < n; i++) {
		v3d p = ldr(pos + 3 * i);
		str(prev + 3 * i, p);
		v3d u1 = sample_v(g, vel, p);
		v3d pm = gc_move(p, u1, 0.5 * dt);
		v3d u2 = vtan(sample_v(g, vel, pm), p);
		str(pos + 3 * i, gc_move(p, u2, dt));
	}
}

/* ---------------- trail renderer (reference src/render.c) ---------------- */
typedef struct {
	int R;
	float *buf;      /* caller-owned [6*R*R*4] */
	double g;
	float radius, alpha;
} ORender;

/* render.c:11-22: radius clamp; buffer starts at 0 and g at 1 */
void orc_render_init(ORender *r, int R, float radius, float alpha, float *buf)
{
	r->R = R;
	r->radius = radius < 0.75f ? 0.75f : radius;
	r->alpha = alpha;
	r->g = 1.0;
	r->buf = buf;
	memset(buf, 0, sizeof(float) * 6 * (size_t) R * R * 4);
}

/* render.c:55-66 */
void orc_render_fade(ORender *r, double fade)
{
	r->g *= (1.0 - fade);
	if (r->g < 1e-4) {
		size_t n = (size_t) 6 * r->R * r->R * 4;
		float g = (float) r->g;
		for (size_t i = 0; i < n; i++)
			r->buf[i] *= g;
		r->g = 1.0;
	}
}

/* One texel-footprint blend of a splat (render.c:113-154 without the strip split: every
 * texel is blended exactly once per record, so splitting a footprint across strips does
 * not change any texel's value). */
static void splat(ORender *r, int f, float px, float py, const float *col, float invg)
{
	int R = r->R;
	float rad = r->radius, invrad = 1.0f / rad;
	float w0 = r->alpha * col[3];
	int ylo = (int) ceilf(py - rad), yhi = (int) floorf(py + rad);
	int xlo = (int) ceilf(px - rad), xhi = (int) floorf(px + rad);
	if (ylo < 0) ylo = 0;
	if (yhi > R - 1) yhi = R - 1;
	if (xlo < 0) xlo = 0;
	if (xhi > R - 1) xhi = R - 1;
	float *face = r->buf + (size_t) f * R * R * 4;
	for (int y = ylo; y <= yhi; y++) {
		float wy = 1.0f - fabsf(y - py) * invrad;
		if (wy <= 0)
			continue;
		float *row = face + (size_t) y * R * 4;
		for (int x = xlo; x <= xhi; x++) {
			float wx = 1.0f - fabsf(x - px) * invrad;
			if (wx <= 0)
				continue;
			float w = w0 * wx * wy;
			float keep = 1.0f - w, add = w * invg;
			float *d = row + 4 * x;
			d[0] = d[0] * keep + add * col[0];
			d[1] = d[1] * keep + add * col[1];
			d[2] = d[2] * keep + add * col[2];
			d[3] = d[3] * keep + add;
		}
	}
}

/* render.c:87-111 (face loop and gnomonic pixel coordinates) + render.c:156-218 (sample
 * placement).  Particles and samples are visited in (particle, sample) order, which is
 * the per-texel blend order the reference's bin-major, thread-minor sort produces. */
void orc_render_deposit(ORender *r, int n, const real *prev, const real *pos, const float *col)
{
	double spacing = 0.7 / r->R;
	int R = r->R;
	float rad = r->radius;
	float invg = (float) (1.0 / r->g);
	for (int i = 0; i < n; i++) {
		v3d a = ldr(prev + 3 * i), b = ldr(pos + 3 * i);
		double ang = vangle(a, b);
		int ns = (int) ceil(ang / spacing);
		if (ns < 1) ns = 1;
		if (ns > 4096) ns = 4096;
		for (int k = 1; k <= ns; k++) {
			double tt = (double) k / ns;
			v3d d = vnorm(vadd(vscl(a, 1 - tt), vscl(b, tt)));
			for (int f = 0; f < 6; f++) {
				double dc = vdot(d, FC[f]);
				if (dc < 0.5)
					continue;
				double u = vdot(d, FR[f]) / dc;
				double v = vdot(d, FU[f]) / dc;
				float px = (float) ((u + 1.0) * 0.5 * R - 0.5);
				float py = (float) ((1.0 - v) * 0.5 * R - 0.5);
				int xlo = (int) ceilf(px - rad), xhi = (int) floorf(px + rad);
				int ylo = (int) ceilf(py - rad), yhi = (int) floorf(py + rad);
				if (xlo < 0) xlo = 0;
				if (xhi > R - 1) xhi = R - 1;
				if (ylo < 0) ylo = 0;
				if (yhi > R - 1) yhi = R - 1;
				if (xlo > xhi || ylo > yhi)
					continue;
				splat(r, f, px, py, col + 4 * (size_t) i, invg);
			}
		}
	}
}

/* Number of splat records the reference emits for these particles (one per face and
 * 32-row strip touched; render.c:23-30, 87-111).  Used for throughput accounting. */
long long orc_render_count_records(const ORender *r, int n, const real *prev, const real *pos)
{
	double spacing = 0.7 / r->R;
	int R = r->R;
	float rad = r->radius;
	int sh = 32;
	if (sh < (int) ceilf(rad) + 1)
		sh = (int) ceilf(rad) + 1;
	if (sh > R)
		sh = R;
	long long cnt = 0;
	for (int i = 0; i < n; i++) {
		v3d a = ldr(prev + 3 * i), b = ldr(pos + 3 * i);
		double ang = vangle(a, b);
		int ns = (int) ceil(ang / spacing);
		if (ns < 1) ns = 1;
		if (ns > 4096) ns = 4096;
		for (int k = 1; k <= ns; k++) {
			double tt = (double) k / ns;
			v3d d = vnorm(vadd(vscl(a, 1 - tt), vscl(b, tt)));
			for (int f = 0; f < 6; f++) {
				double dc = vdot(d, FC[f]);
				if (dc < 0.5)
					continue;
				double u = vdot(d, FR[f]) / dc;
				double v = vdot(d, FU[f]) / dc;
				float px = (float) ((u + 1.0) * 0.5 * R - 0.5);
				float py = (float) ((1.0 - v) * 0.5 * R - 0.5);
				int xlo = (int) ceilf(px - rad), xhi = (int) floorf(px + rad);
				int ylo = (int) ceilf(py - rad), yhi = (int) floorf(py + rad);
				if (xlo < 0) xlo = 0;
				if (xhi > R - 1) xhi = R - 1;
				if (ylo < 0) ylo = 0;
				if (yhi > R - 1) yhi = R - 1;
				if (xlo > xhi || ylo > yhi)
					continue;
				cnt += yhi / sh - ylo / sh + 1;
			}
		}
	}
	return cnt;
}

double orc_render_g(const ORender *r) { return r->g; }

/* render.c:220-226 */
void orc_render_texel(const ORender *r, int face, int x, int y, float out[4])
{
	const float *d = r->buf + (((size_t) face * r->R + y) * r->R + x) * 4;
	float g = (float) r->g;
	for (int c = 0; c < 4; c++)
		out[c] = d[c] * g;
}

/* render.c:253-258 */
static inline unsigned char to8(float v)
{
	if (v <= 0) return 0;
	if (v >= 1) return 255;
	return (unsigned char) (v * 255.0f + 0.5f);
}

/* render.c:260-279 */
void orc_render_pixel8(const float c[4], const float *bg, unsigned char *out)
{
	float a = c[3];
	if (a > 1) a = 1;
	if (a < 0) a = 0;
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

/* 8-bit faces of the whole buffer, texel by texel as render_texel then render_pixel8:
 * [6][R][R][4] un-premultiplied RGBA when bg is NULL, else [6][R][R][3] composited over
 * bg (the per-face loop of main.c:258-276). */
void orc_render_faces8(const ORender *r, const float *bg, unsigned char *out)
{
	int R = r->R, ch = bg ? 3 : 4;
	#pragma omp parallel for schedule(static)
	for (int f = 0; f < 6; f++)
		for (int y = 0; y < R; y++)
			for (int x = 0; x < R; x++) {
				float c[4];
				orc_render_texel(r, f, x, y, c);
				orc_render_pixel8(c, bg, out + ((((size_t) f * R + y) * R) + x) * ch);
			}
}

/* render.c:228-251: bilinear sample of the faces in direction d */
void orc_render_sample_dir(const ORender *r, double dx, double dy, double dz, float out[4])
{
	v3d d = mk(dx, dy, dz);
	int f = face_of(d);
	double dc = vdot(d, FC[f]);
	double u = vdot(d, FR[f]) / dc;
	double v = vdot(d, FU[f]) / dc;
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
	orc_render_texel(r, f, x0, y0, a);
	orc_render_texel(r, f, x1, y0, b);
	orc_render_texel(r, f, x0, y1, c);
	orc_render_texel(r, f, x1, y1, e);
	for (int k = 0; k < 4; k++)
		out[k] = (1 - fy) * ((1 - fx) * a[k] + fx * b[k]) + fy * ((1 - fx) * c[k] + fx * e[k]);
}

/* main.c:280-295: the 4R x 3R cross net, RGBA, cells (0,1) (1,1) (2,1) (3,1) (0,0) (0,2)
 * for faces 0-5, uncovered pixels transparent black; a face pixel's alpha byte is 255
 * when bg is given (render_pixel8 leaves it). */
void orc_render_net8(const ORender *r, const float *bg, unsigned char *out)
{
	static const int cell_x[6] = { 0, 1, 2, 3, 0, 0 }, cell_y[6] = { 1, 1, 1, 1, 0, 2 };
	int R = r->R, W = 4 * R;
	memset(out, 0, (size_t) W * 3 * R * 4);
	#pragma omp parallel for schedule(static)
	for (int fy = 0; fy < 6 * R; fy++) {
		int f = fy / R, y = fy % R;
		for (int x = 0; x < R; x++) {
			float c[4];
			unsigned char px[4] = { 0, 0, 0, 255 };
			orc_render_texel(r, f, x, y, c);
			orc_render_pixel8(c, bg, px);
			memcpy(out + (((size_t) cell_y[f] * R + y) * W + cell_x[f] * R + x) * 4, px, 4);
		}
	}
}

/* main.c:297-312 with vec3.h:55-58: the equirectangular preview, W x H pixels (the
 * reference uses W = 2R, H = R), 4 channels without bg, 3 with. */
void orc_render_preview8(const ORender *r, int W, int H, const float *bg, unsigned char *out)
{
	int ch = bg ? 3 : 4;
	#pragma omp parallel for schedule(static)
	for (int y = 0; y < H; y++) {
		double lat = M_PI / 2 - (y + 0.5) / H * M_PI;
		for (int x = 0; x < W; x++) {
			double lon = (x + 0.5) / W * 2 * M_PI - M_PI;
			float c[4];
			orc_render_sample_dir(r, cos(lat) * sin(lon), sin(lat), cos(lat) * cos(lon), c);
			orc_render_pixel8(c, bg, out + ((size_t) y * W + x) * ch);
		}
	}
}
