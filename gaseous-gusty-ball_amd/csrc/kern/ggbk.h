/*
 * ggbk.h -- the thin C ABI between the host C code and the gfx950 HIP kernels.
 *
 * Plain C structs of device pointers and sizes, passed by value; every launcher
 * enqueues on the library's one stream and returns immediately unless its name says
 * it synchronises.  Device arrays use "column-major ELL" for the Voronoi stencil:
 * slot k of cell i lives at [k * n + i], so the threads of a wave read consecutive
 * words for every slot.
 *
 * Precision codes: the state type T and the multigrid type TM are selected per call by
 * an int (GGBK_F64 / GGBK_F32).
 */
#ifndef GGBK_H
#define GGBK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GGBK_F64 = 0, GGBK_F32 = 1 };

/* one grid level on the device */
typedef struct {
	int N, n, maxnb, pad_;
	double h, total_area;
	const float *pos;         /* [3n] AoS cell centres */
	const float *area;        /* [n] */
	const unsigned char *nnb; /* [n] */
	const int *nbr;           /* [maxnb][n] */
	const float *w;           /* [maxnb][n] */
	const float *elen;        /* [maxnb][n] */
	const float *gc;          /* [maxnb][3][n] */
	const int *ghost;         /* [6][N+2][N+2] tap index of (f, i, j), i, j in [-1, N] */
	const float *dinv;        /* [n] 1 / sum(w), rounded to float */
	/* compressed stencil: per cell (class | irregular << 31, offset codes or explicit row) */
	const int *srec;          /* [n][2] */
	const float *scw;         /* [ncls][8] weights of each class, slot order, zero-padded */
	const int *sxn;           /* [nirr][8] neighbours of the irregular cells, self-padded */
	int ncls, nirr;
} ggbk_grid;

/* bilinear prolongation from level l+1 to l, and its transpose.  Both are stored per
 * row as interleaved (index, weight bits) pairs, so a row is read with 16-byte loads. */
typedef struct {
	int nf, nc, rmax, rpad;  /* rpad: padded row width, 18 or GGBK_RMAX */
	int Nc, analytic;  /* coarse face size; indices of face-interior rows follow the
	                    * bilinear pattern (only the weights are read for those) */
	int rana;          /* the pattern holds (rw16 is then tabulated), whatever the size */
	int exact;         /* the pattern holds and the face-interior weights are the exact
	                    * bilinear ones, products of 1/4 and 3/4 by the cells' parity: they are
	                    * computed, and pw / rw16 are not stored (NULL) */
	const float *pw;   /* [nf][4] weights of each fine cell's taps */
	const float *rw16; /* [4][nc][4] weights of the 16 entries of face-interior rows of R, in
	                    * row order (entry e in element e / 4, lane e % 4) */
	const int *pt;     /* [nf][4][2] coarse tap and weight of each fine cell */
	const int *rr;     /* [rpad/2][nc][4] rows of R = P^T, two (fine index, weight bits) pairs
	                    * per element, column-major over rows (coalesced across a wave); fine
	                    * index ascending, zero-weight padded */
} ggbk_xfer;
#define GGBK_RMAX 24

/* Device-resident solver scalars.  Reductions write per-block partials; the last block
 * to finish folds them in block order (deterministic) and updates these. */
typedef struct {
	double bb, rr, rz, pAp, alpha, beta, tol2;
	double rznew;
	double sum;        /* result of the last generic sum */
	double vmax2;      /* max |vel|^2 of the last velocity pass */
	double mean;       /* area-weighted mean of the last remove-mean */
	int it, maxit;
	int active;        /* CG loop running (the next update is applied) */
	int status;        /* 0 running/converged, 1 pAp <= 0, 2 bb == 0, 3 maxit */
	int vc;            /* mixed-precision PCG: the passes after the next update are needed */
	int stop;          /* mixed-precision PCG: the next update is the last */
	unsigned int ticket[8];
} ggbk_scalars;

enum { GGBK_TICKETS = 8 };

/* ---- runtime ---- */
int ggbk_init(void);                       /* select device 0, create the stream; 0 = ok */
const char *ggbk_last_error(void);
void *ggbk_malloc(size_t bytes);           /* NULL on failure */
void ggbk_free(void *p);
int ggbk_h2d(void *dst, const void *src, size_t bytes);      /* synchronous */
int ggbk_d2h(void *dst, const void *src, size_t bytes);      /* synchronous */
int ggbk_d2h_async(void *dst, const void *src, size_t bytes); /* ordered on the stream */
int ggbk_d2d(void *dst, const void *src, size_t bytes);
int ggbk_zero(void *p, size_t bytes);      /* fill kernel (graph-safe, no memset node) */
int ggbk_sync(void);
void *ggbk_host_alloc(size_t bytes);       /* pinned */
void ggbk_host_free(void *p);
int ggbk_check(const char *where);         /* reports a pending launch error */
/* stream timing */
void *ggbk_event_new(void);
void ggbk_event_record(void *ev);
float ggbk_event_ms(void *a, void *b);     /* synchronises on b */
void ggbk_event_free(void *ev);
void *ggbk_stream_handle(void);

/* ---- launch probes (HIP events around selected kernels, for measurement) ---- */
enum {
	GGBK_PROBE_RENDER_DRAW = 1,
	GGBK_PROBE_RENDER_COUNT = 2, /* binning: count + scan */
	GGBK_PROBE_MG_UP0 = 3,       /* level-0 prolongation + post-smooth */
	GGBK_PROBE_MG_DOWN0 = 4,     /* level-0 pre-smooth residual */
	GGBK_PROBE_CG_AP = 5,
	GGBK_PROBE_CG_UPDATE = 6,
	GGBK_PROBE_PARTICLES = 7,
	GGBK_PROBE_DEPARTURES = 8,
	GGBK_PROBE_RENDER_FILL = 9,  /* binning: list fill + reordered particle copy */
};
void ggbk_probe_enable(int id, int on);
int ggbk_probe_collect(int id, double *total_ms);
void ggbk_copy_kernel(void *dst, const void *src, size_t bytes);
#ifdef __cplusplus
int ggbk_probe_begin(int id);
void ggbk_probe_end(int id);
#endif

/* ---- conversions ---- */
void ggbk_convert(void *dst, int tdst, const void *src, int tsrc, size_t count);

/* ---- free-standing operators (reference fluid.c:211-275) ---- */
void ggbk_curl(ggbk_grid g, const void *vel, void *out, int T);
void ggbk_divergence(ggbk_grid g, const void *vel, void *out, int T);
void ggbk_laplacian(ggbk_grid g, const void *x, void *y, int T);
void ggbk_applyK(ggbk_grid g, const void *x, double *y, int T);

/* ---- reductions ----
 * Reduction kernels store one partial per block in slot `slot` of a partials array
 * (slot stride ggbk_red.stride doubles); a one-workgroup fold kernel then combines the
 * partials in block order and updates ggbk_scalars.  Deterministic.  The mixed-precision
 * PCG's fold of nine sums runs as nine workgroups, one per sum (into slot G0 + q); the
 * last to finish (a ticket counter in cnt, no waiting) updates the scalars. */
typedef struct {
	double *part;
	int stride;
	unsigned *cnt;  /* ticket counters (zero between launches), or NULL */
} ggbk_red;
enum { GGBK_SL_BB = 0, GGBK_SL_RR0, GGBK_SL_PAP, GGBK_SL_RR, GGBK_SL_RZ, GGBK_SL_MEAN, GGBK_SL_VMAX,
       /* the mixed-precision PCG's fold after each V-cycle (z its output, r the residual it
        * preconditioned, Ap the current search direction's image) */
       GGBK_SL_ZAZ, GGBK_SL_ZAP, GGBK_SL_RAZ, GGBK_SL_AZAZ, GGBK_SL_AZAP, GGBK_SL_RAP, GGBK_SL_APAP,
       /* the nine sums RZ, ZAZ .. APAP, RR of its fold, in that order (element 0 of each) */
       GGBK_SL_G0, GGBK_NSLOTS = GGBK_SL_G0 + 11 };
#define GGBK_NCNT 8 /* ticket counters per ggbk_red */
enum {
	GGBK_FIN_START = 0, /* bb, rr -> tol2, active, status */
	GGBK_FIN_PAP,       /* pAp -> alpha, or stop */
	GGBK_FIN_RR,        /* rr -> it++, stop on convergence or maxit */
	GGBK_FIN_RZ_FIRST,  /* rz */
	GGBK_FIN_RZ,        /* rznew -> beta */
	GGBK_FIN_MEAN_F64,  /* mean = sum / total_area */
	GGBK_FIN_MEAN_F32,  /* mean rounded to float */
	GGBK_FIN_VMAX,      /* vmax2 = max */
	GGBK_FIN_CHEB_END,  /* mixed-precision PCG, after an update applied alone: the loop ends
	                     * if it was the last */
};
/* nb = number of blocks that wrote partials */
void ggbk_fin(int op, ggbk_scalars *s, ggbk_red red, int nb, double tol, int maxit, double total_area);
int ggbk_nblocks(int n);

/* f's area-weighted sum partials (slot MEAN); fold with FIN_MEAN_* */
void ggbk_mean_partial(ggbk_grid g, const void *f, int T, ggbk_red red);
/* after a solve: psi = 0 when s->status says b was zero, then psi's mean partial */
void ggbk_mean_partial_bb0(ggbk_grid g, void *psi, int T, const ggbk_scalars *s, ggbk_red red);
void ggbk_sub_mean(int n, void *f, int T, const ggbk_scalars *s);
/* f -= area-weighted mean, rounded to T as the reference rounds it (fluid.c:264-275) */
void ggbk_remove_mean(ggbk_grid g, void *f, int T, ggbk_scalars *s, ggbk_red red);
/* The same on float f in device memory, with the device's static scratch in place of
 * allocated scalars and partials: 0, -1 (scratch not set up), -2 (more than
 * GGBK_NOALLOC_CELLS cells).  ggbk_noalloc_init looks the scratch up on the current device. */
#define GGBK_NOALLOC_CELLS 6291456
int ggbk_noalloc_init(void);
int ggbk_remove_mean_noalloc(ggbk_grid g, float *f);

/* ---- PCG (reference fluid.c:288-358) ---- */
/* [omega -= mean if sub_mean] r = b - K psi with b = -area*omega; bb, rr partials; with
 * open_vc it sets s->vc, so that the mixed PCG's first V-cycle runs before its fold starts
 * the solve */
void ggbk_cg_start(ggbk_grid g, void *omega, const void *psi, void *r, int T, int sub_mean,
		   ggbk_scalars *s, ggbk_red red, int open_vc);
/* p_out = z (first) or z + beta p_in; Ap = K p_out; pAp partial (guarded by active) */
void ggbk_cg_Ap(ggbk_grid g, const void *z, int TZ, const void *p_in, void *p_out, void *Ap, int T,
		int first, const ggbk_scalars *s, ggbk_red red);
/* psi += alpha p; r -= alpha Ap; rr partial */
void ggbk_cg_update(ggbk_grid g, void *psi, void *r, const void *p, const void *Ap, int T,
		    const ggbk_scalars *s, ggbk_red red);
/* Jacobi preconditioner z = dinv r and the r.z partial (no multigrid) */
void ggbk_jacobi_precond(ggbk_grid g, const void *r, void *z, int T, const ggbk_scalars *s, ggbk_red red);
/* bb == 0 exit: psi = 0 when status says so */
void ggbk_cg_zero_if_bb0(ggbk_grid g, void *psi, int T, const ggbk_scalars *s);

/* ---- multigrid V-cycle (reference fluid.c:100-146) ---- */
/* Level-0 input r (type T) -> output z (type TM), intermediate arrays of type TM.
 * levels[l].x/b/r are the per-level TM scratch; level 0's x and b are unused.  A level
 * with M != NULL applies the precomputed matrix of the V-cycle below it (x = M b;
 * M is [n][n] row-major, double when TM is double else float). */
/* 16x16 face tiles of a level for the fused multi-sweep smoothing passes (mixed precision):
 * a workgroup owns a tile (ring 0) and recomputes its `deg` surrounding rings of cells, so
 * `deg` Chebyshev steps need no other synchronisation than the workgroup's own barriers.
 * An interior tile (kind -1) has its rings on its own face with regular stencils, found by
 * index arithmetic; a tile near a face seam has graph tables: its ring cells in BFS order
 * over the Voronoi stencil, and each stencil neighbour as a slot of that list. */
typedef struct {
	int ntiles, deg, cap, ngraph;   /* cap: largest ext cell count of a graph tile */
	int geo;                        /* 1: the interior tiles are exactly those whose hs
	                                 * rings lie within their face (the kernel then tells
	                                 * them by their index alone) */
	int hs;                         /* rings of an interior tile */
	int nslot;                      /* stencil slots per ext cell in gnbr: 8 (Voronoi) or
	                                 * GGBK_WNB (a Galerkin level) */
	int pad_;
	const int *tinfo;               /* [ntiles][GGBK_TINFO] kind (-1 interior, else graph
	                                 * index), first ext cell, ext count, ext cells in rings
	                                 * 0..r (r <= 5) */
	const int *gcc;                 /* [ext][2] cell index and stencil weight class of each
	                                 * ext cell of the graph tiles, ring by ring */
	const unsigned short *gnbr;     /* [ext][nslot] slot of each stencil neighbour, self-padded */
	/* tiles of a down pass with the restriction fused: ring 0 of a graph tile is the fine
	 * support of the rows of R of the tile's 8x8 coarse cells (its own 16x16 cells first),
	 * and each row is listed by slot; an interior tile's rows follow the bilinear pattern */
	const unsigned short *grow;     /* [ngraph][64][GGBK_TROW] slots of each coarse row */
	const float *grw;               /* [ngraph][64][GGBK_TROW] their weights (zero-padded) */
} ggbk_tiles;
#define GGBK_TROW GGBK_RMAX
#define GGBK_TINFO 12

/* A Galerkin coarse operator of the mixed-precision hierarchy, K_c = R K_f P, in the
 * form (K x)_i = sum_k w_ik (x_i - x_nb(k)) with w_ik = -(R K_f P)_ij (zero row sums).  A
 * regular cell (its neighbours on its own face within two cells each way) has its slots
 * at the fixed offsets of the 5x5 square around it, row-major without the centre; any
 * other cell lists its neighbours by ascending index (in the graph tiles' tables).
 * Weights are stored as IEEE half floats. */
#define GGBK_WNB 24
typedef struct {
	int n, N;
	const unsigned short *w;  /* [n][GGBK_WNB] f16 bits of w_ik, slot order, zero-padded */
	const float *dinv;        /* [n] 1 / sum_k w_ik */
} ggbk_wide;

typedef struct {
	ggbk_grid g;
	ggbk_wide wd;     /* the Galerkin operator of a coarse level of the mixed hierarchy
	                   * (wd.w == NULL: the level's own Voronoi stencil) */
	ggbk_xfer xf;     /* transfer to the next coarser level (unused on the coarsest) */
	void *x, *b, *r;  /* TM scratch */
	const void *M;    /* dense sub-V-cycle operator, or NULL */
	ggbk_tiles tl;    /* multi-sweep tiles of the up pass (tl.ntiles == 0: none) */
	ggbk_tiles td;    /* tiles of the down pass (restriction into the next level fused) */
	void *xpre;       /* TM: the pre-smoothed x of the tiled passes */
} ggbk_level;
/* z = V(r); with red.part != NULL also the r.z partial (slot RZ) of ggbk_nblocks(n0) blocks */
void ggbk_vcycle(const ggbk_level *levels, int nlev, const void *rin, int T, void *zout, int TM,
		 ggbk_scalars *s, ggbk_red red);
/* PCG update fused into the V-cycle's first pass: psi += alpha p; r_out = r_in - alpha Ap
 * (fold FIN_RR enqueued after it); then z = V(r_out) as ggbk_vcycle.  Needs nlev >= 2;
 * returns -1 (nothing enqueued) otherwise.  r_out must not alias r_in. */
typedef struct {
	void *psi;
	const void *r_in, *p, *Ap;
} ggbk_cg_fuse;
int ggbk_vcycle_update(const ggbk_level *levels, int nlev, const ggbk_cg_fuse *u, void *r_out, int T,
		       void *zout, int TM, ggbk_scalars *s, ggbk_red red);
/* The mixed-precision preconditioner: a V-cycle with Chebyshev smoothing (level 0 of
 * degree levels[0].td.deg, 2-4; the coarse levels of degree 2 on their Galerkin operators
 * levels[l].wd, or of degree 3 on their own Voronoi operators; 1/2 <= lambda(D^-1 K) <= 2,
 * the Gershgorin bound) on the tiled levels, float storage and arithmetic, down to the
 * first level with a dense operator (levels[0..L-1] tiled, levels[L].M set, L >= 1; level
 * 0's up tiles have one ring more than its degree).
 * c == NULL: z = M r (stores gated by s->vc).  Otherwise the PCG's V-cycle: with c->update
 * its first pass applies the update (the direction p = z + beta p, Ap = Az + beta Ap, first:
 * p = z, Ap = Az; psi += alpha p; r_out = r - alpha Ap) and preconditions r_out, else it
 * preconditions c->r; its last pass also stores A z (double) and the partials of the
 * sums of ggbk_cheb_fold, which ggbk_fold_cheb then folds into the scalars.  The update's
 * stores are gated by s->active, the V-cycle's by s->vc.  r_out must not
 * alias r, nor Ap_out Ap_in. */
typedef struct {
	int update;                   /* apply the PCG update in the first pass */
	int first;                    /* the update is the first: p = z, Ap = Az */
	double *psi;
	const double *r;              /* the residual before the update (update) or to precondition */
	double *r_out;                /* update: r - alpha Ap */
	const double *p_in;           /* the previous direction (may alias p_out) */
	double *p_out;
	const double *Ap_in;          /* A p of the previous direction (NULL before the first) */
	double *Ap_out;
	double *Az;                   /* in: the previous V-cycle's A z (update); out: this one's */
	const float *z;               /* the previous V-cycle's z (update) */
	int start;                    /* the solve's first V-cycle, whose fold also starts the solve */
	int nb0;                      /* start: ggbk_cg_start's partials per sum */
	int maxit, pad_;              /* start: the iteration cap */
	double tol;                   /* start: the relative tolerance */
} ggbk_cg_cheb;
/* The fold after the PCG's V-cycle (nine workgroups; nb partials per sum): ggbk_cheb_fold
 * (dev.h) on the nine sums.  first: the solve's first V-cycle (no direction yet).  With
 * nb0 > 0 it also starts the solve from ggbk_cg_start's bb and rr partials (nb0 each, two
 * more workgroups), as ggbk_fin GGBK_FIN_START would, before that fold. */
void ggbk_fold_cheb(ggbk_scalars *s, ggbk_red red, int nb, int first, int nb0, double tol, int maxit);
/* The PCG update alone, as ggbk_vcycle_cheb's first pass applies it: for the update the
 * host expects to be the last of the solve. */
void ggbk_cg_update_cheb(ggbk_grid g, const ggbk_cg_cheb *c, const ggbk_scalars *s);
void ggbk_vcycle_cheb(const ggbk_level *levels, int nlev, const ggbk_cg_cheb *c, const void *r, void *zout,
		      ggbk_scalars *s, ggbk_red red);

/* ---- velocity (reference fluid.c:360-376) + max speed (fluid.c:480-491) ---- */
/* [psi_out = psi_in - mean] vel = p x grad(psi_out); vmax partial (slot VMAX) */
void ggbk_velocity(ggbk_grid g, const void *psi_in, void *psi_out, int sub_mean, void *vel, int T,
		   const ggbk_scalars *s, ggbk_red red);
void ggbk_max_speed2(ggbk_grid g, const void *vel, int T, ggbk_red red);

/* ---- semi-Lagrangian step (reference fluid.c:397-478) ---- */
/* departure points and q0 = omega + 2 Omega y */
void ggbk_departures(ggbk_grid g, const void *vel, const void *omega, void *q0, void *dep_b, void *dep_f,
		     int T, double dt, double twoOmega, int fwd);
void ggbk_advect(ggbk_grid g, const void *src, const void *dep, void *dst, int T);
/* q2 = q + 0.5 (q - A_f q1) */
void ggbk_advect_bfecc(ggbk_grid g, const void *q1, const void *q, const void *dep_f, void *q2, int T);
/* qn = limited A_b q2 (bfecc) or qn = q1 (plain); omega = qn - 2 Omega y relaxed to band;
 * area-weighted omega partial (slot MEAN) */
void ggbk_limited_relax(ggbk_grid g, const void *q2, const void *q, const void *dep_b, const void *q1, int bfecc,
			const void *band, void *omega, int T, double twoOmega, double relax, ggbk_red red);
/* the same with the solve's warm start by extrapolation in time in the same pass (FP64
 * state): psi <- c0 psi + c1 h1 + c2 h2, h2 <- the old psi */
void ggbk_limited_relax_extrap(ggbk_grid g, const double *q2, const double *q, const double *dep_b, const double *q1,
			       int bfecc, const double *band, double *omega, double twoOmega, double relax, ggbk_red red,
			       double *psi, const double *h1, double *h2, double c0, double c1, double c2);

/* ---- discrete adjoint of the fluid step (FP64 state; adj_k.hip) ---- */
void ggbk_adj_sum(ggbk_grid g, const double *x, ggbk_red red);     /* slot MEAN partials of sum x */
/* The transposes of the step's interpolations, without floating-point atomics: the
 * scatter passes below write per-source entries (target index or -1, W values), and
 * ggbk_adj_transpose adds to every target the sum of its entries in entry (source) order.
 * Buffers for up to cap_entries entries and cap_targets targets. */
typedef struct {
	int *tgt;         /* [cap_entries] */
	double *val;      /* [cap_entries][3] */
	int *list;        /* [cap_entries] entry ids by target */
	int *cnt, *cursor; /* [cap_targets] */
	int *start;       /* [cap_targets + 1] */
	int *bsum;        /* [cap_targets / 1024 + 1] */
	long long cap_entries;
	int cap_targets, pad_;
} ggbk_adj_tr;
/* d0[t] (t < split) or d1[t - split] (W values each) += the entries of target t in order;
 * nent entries, nt targets.  -1 when the sizes exceed the buffers. */
int ggbk_adj_transpose(const ggbk_adj_tr *t, long long nent, int W, int nt, int split, double *d0, double *d1);
/* 4 entries per cell: targets [0, n) for lq, [n, 2n) for lq2 (W = 1) */
void ggbk_adj_limiter(ggbk_grid g, const double *q, const double *q2, const double *dep_b,
		      const double *band, const double *lam, const ggbk_scalars *s, double relax,
		      double drelax, double c, int bfecc, const ggbk_adj_tr *tr, double *ldb, ggbk_red red);
/* lq += 1.5 lq2; 4 entries per cell for lq1 (W = 1) */
void ggbk_adj_bfecc_f(ggbk_grid g, const double *lq2, const double *q1, const double *dep_f, double *lq,
		      const ggbk_adj_tr *tr, double *ldf);
/* 4 entries per cell for lq (W = 1) */
void ggbk_adj_advect_b(ggbk_grid g, const double *lq1, const double *q, const double *dep_b,
		       const ggbk_adj_tr *tr, double *ldb);
/* 10 entries per cell for lvel (W = 3) */
void ggbk_adj_departures(ggbk_grid g, const double *vel, const double *ldb, const double *ldf, double dt,
			 int bfecc, const ggbk_adj_tr *tr);
void ggbk_adj_velocity(ggbk_grid g, const double *lvel, const unsigned char *rslot, double *lpsi, ggbk_red red);
void ggbk_adj_rhs(ggbk_grid g, const double *lpsi, const ggbk_scalars *s, double sum_area, double *om);
void ggbk_adj_combine(ggbk_grid g, const double *lq, const double *z, double *lam);
void ggbk_adj_project(ggbk_grid g, const double *lam, const ggbk_scalars *s, double *out);
void ggbk_adj_accum(const ggbk_scalars *s, double *acc);

/* ---- particles (reference particles.c) ----
 * Particle arrays in storage order: slot s holds particle pid[s].  The renderer keeps
 * the slots in tile order (each deposit writes the next order into a second set). */
typedef struct {
	int n, pad_;
	double *pos, *prev;   /* [3n] */
	float *col;           /* [4n] */
	int *pid;             /* [n] */
} ggbk_parts;
void ggbk_particles_init(ggbk_parts p, int round32);  /* identity order */
void ggbk_particles_color(int n, const double *pos, float *col, const float *img, int w, int h);
void ggbk_particles_advect(ggbk_grid g, int np, double *pos, double *prev, const void *vel, int T,
			   double dt, int round32);

/* ---- renderer (reference render.c) ---- */
typedef struct {
	int R, tile, ntx, ntiles;   /* tile edge in texels, tiles per face edge, 6*ntx*ntx */
	float radius, alpha;
	float *buf;                 /* [6][R][R][4] */
	int *tile_count;            /* [ntiles + 1] */
	int *tile_start;            /* [ntiles + 1] */
	int *tile_cursor;           /* [ntiles] */
	int *home_count;            /* [ntiles + 1] particles whose current position is on the tile */
	int *home_start;            /* [ntiles + 1] */
	int *home_cursor;           /* [ntiles] */
	int *list;                  /* [cap] particle slots per tile */
	int *bucket;                /* [ntiles][tile_bucket()] particle slots, while no tile has more */
	int bucket_mode;            /* the draw reads the buckets (1) or the list (0) */
	unsigned short *nsamp;      /* [nsamp_cap] sample count of each particle slot (filled by the binning) */
	long long nsamp_cap;
	long long cap;
	unsigned long long *sort_scratch; /* [scratch_cap] for tiles whose list exceeds the LDS sort */
	long long scratch_cap;
	long long *total;           /* [16] entries, longest list, draw diagnostics (device) */
	int *tile_diag;             /* [8 * ntiles] per-tile draw diagnostics */
	long long *scan_part;       /* [3 * ggbk_render_scan_blocks(ntiles)] */
	long long *hstats;          /* [2] pinned host copy of total[0..1] */
	void *stats_ev;             /* event recorded after that copy */
	void *fill_ev;              /* event recorded after the binning */
	void *copy_stream;          /* stream of that copy (default priority) */
} ggbk_render;
int ggbk_render_host_init(ggbk_render *r); /* hstats, the events, the copy stream */
void ggbk_render_host_free(ggbk_render *r);
int ggbk_render_tile(void);     /* tile edge the kernels are built for */
int ggbk_render_scan_blocks(int ntiles);
void ggbk_render_scale(float *buf, size_t count, float g);
/* 8-bit texels of the whole buffer scaled by g (the reference's render_texel + to8 /
 * render_pixel8): ch = 4 un-premultiplied RGBA (bg NULL), ch = 3 composited over bg */
void ggbk_render_faces8(const float *buf, size_t ntexels, float g, const float *bg, unsigned char *out);
void ggbk_render_net8(const float *buf, int R, float g, const float *bg, unsigned char *out);
void ggbk_render_preview8(const float *buf, int R, float g, int W, int H, const float *bg, unsigned char *out);
/* sums the per-tile diagnostics of the last deposit and copies the 16 values
 * (synchronises) */
int ggbk_render_stats(const ggbk_render *r, long long *out16);
int ggbk_render_bucket(void);  /* slots per tile bucket */
/* bin + ordered blend of the particles `in`; with reorder != 0 also writes them to `out`
 * in the next storage (tile) order.  Returns the number of (particle, tile) entries, or
 * -1 on error. */
long long ggbk_render_deposit(ggbk_render *r, ggbk_parts in, ggbk_parts out, float invg, int reorder);

#ifdef __cplusplus
}
#endif

#endif
