/* Extensions to the reference API (no counterpart in the reference headers).
 *
 * The reference keeps all state in host float arrays; this library keeps it on the GPU
 * (in double by default).  These calls move state between the device and the host
 * mirrors, expose the full-precision state for verification, report device timings,
 * and carry the discrete adjoint.
 */
#ifndef GGB_EXT_H
#define GGB_EXT_H

#include "fluid.h"
#include "particles.h"
#include "render.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fields addressed by fluid_get_f64 / fluid_set_f64. */
enum {
	GGB_FIELD_OMEGA = 0,      /* [n] */
	GGB_FIELD_OMEGA_BAND = 1, /* [n] */
	GGB_FIELD_PSI = 2,        /* [n] */
	GGB_FIELD_VEL = 3,        /* [3n] */
	GGB_FIELD_DEP_B = 4,      /* [3n] backward departure points of the last step */
	GGB_FIELD_DEP_F = 5,      /* [3n] forward departure points of the last step */
};

/* Selects the GPU and creates the library stream; 0 on success.  Called implicitly by
 * every *_init; fails (-1) when no GPU is present. */
int ggb_device_init(void);
/* Library version string. */
const char *ggb_version(void);
/* Wait for all queued device work. */
void ggb_sync(void);
/* on != 0: fluid_step and particles_advect also refresh the host mirrors (F.omega,
 * F.omega_band, F.psi, F.vel; P.pos, P.prev), as fluid_sync_host / particles_sync_host
 * would, so callers that read them after every step -- the reference's own
 * tests/test_flow.c -- work unchanged, at the cost of a device-to-host copy per call.
 * The initial setting is GGB_SYNC_MIRRORS from the environment (1 = on, default off). */
void ggb_set_sync_mirrors(int on);

/* ---- fluid ---- */
void fluid_sync_host(Fluid *F);   /* device state -> F.omega/omega_band/psi/vel mirrors */
void fluid_sync_device(Fluid *F); /* mirrors -> device state (rounds FP64 state to float) */
int fluid_get_f64(const Fluid *F, int field, double *out);
int fluid_set_f64(Fluid *F, int field, const double *in);
/* PCG on the device state (no mirror traffic); returns iterations. */
int fluid_solve_psi_device(Fluid *F);
/* True relative residual ||b - K psi|| / ||b|| (b = -A omega) of the device state, in FP64. */
double fluid_true_relres(Fluid *F);
/* V-cycle preconditioner applied to b (double[n]) -> x (double[n]); for verification. */
int fluid_vcycle_f64(Fluid *F, const double *b, double *x);

/* ---- free operators in double (device), for verification ---- */
int ggb_laplacian_f64(const CubeSphere *cs, const double *x, double *y);
int ggb_applyK_f64(const CubeSphere *cs, const double *x, double *y);
int ggb_curl_f64(const CubeSphere *cs, const double *vel, double *out);
int ggb_divergence_f64(const CubeSphere *cs, const double *vel, double *out);
/* The compressed stencil the device passes read (host only, no device needed): per cell
 * rec[2i] = weight class (| 1 << 31 for a cell with an explicit neighbour row) and
 * rec[2i+1] = eight 4-bit neighbour codes, slot k in bits 4k..4k+3, code (dy + 1) << 2 |
 * (dx + 1) for neighbour i + dy N + dx (self = 5 pads), or the explicit row;
 * cw[8 * class + k] = weight of slot k
 * (zero-padded); xn[8 * row + k] = neighbour of slot k (self-padded).  Capacities: rec 2n,
 * cw 8n, xn 8n.  Has no counterpart in the reference, whose stencil is the ELL table
 * itself (cubesphere.h nbr/w). */
int ggb_stencil_compress(const CubeSphere *cs, int *rec, float *cw, int *xn, int *ncls, int *nirr);

/* fluid_remove_mean (reference fluid.c:264-275; float f -= its area-weighted mean) on
 * f_dev, n = cs->ncells floats in device memory, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its reduction scratch (the
 * scalar block and per-block partials) is a static array each GPU holds for the life of
 * the process, sized for grids of up to 6,291,456 cells (N = 1024): larger grids return -2
 * with a message.  The grid's device tables must already exist (any earlier operator,
 * fluid_init or particles_advect on this CubeSphere uploads them); otherwise -1.  The work
 * is enqueued on the library stream and not waited for (ggb_sync waits); calls that
 * overlap on one GPU share its scratch and race.  The result equals fluid_remove_mean's
 * bit for bit.  0 on success. */
int fluid_remove_mean_noalloc(const CubeSphere *cs, float *f_dev);

/* ---- particles ---- */
void particles_sync_host(Particles *P); /* device -> pos/prev mirrors */
int particles_get_f64(const Particles *P, double *pos, double *prev);
int particles_set_f64(Particles *P, const double *pos, const double *prev);
/* 1 = round stored positions to float (the reference's storage); 0 = double (default) */
void particles_set_round32(Particles *P, int on);

/* ---- render ---- */
void render_sync_host(Render *r);             /* device buffer -> r->buf mirror */
long long render_last_entries(const Render *r); /* (particle, tile) entries of the last deposit */
/* diagnostics of the last deposit (16 values): [0] (particle, tile) entries, [1] longest
 * tile list, [2] sample ranges, [3] serial ranges, [4] tiles sorted by particle index,
 * [5..7] draw-pass cycle sums (samples, blend, total), [8] texels touched, [9] cycles of
 * the window set-up (part of [5]) */
int render_last_stats(const Render *r, long long *out16);
/* 8-bit faces [6][R][R][4] (bg NULL: un-premultiplied RGBA) or [6][R][R][3] (composited
 * over bg), texel by texel as render_texel then render_pixel8 (reference render.c:220-226,
 * 253-279), converted on the device; only the 8-bit result is copied back. */
void render_faces8(Render *r, const float *bg, unsigned char *out);
/* The 4R x 3R cross net [3R][4R][4] of the reference driver's --net output
 * (main.c:280-295): faces 0-3 along the middle row, 4 above and 5 below face 0, each
 * pixel as render_pixel8 of render_texel (alpha 255 when bg is given); the rest of the
 * net transparent black.  Converted on the device. */
void render_net8(Render *r, const float *bg, unsigned char *out);
/* The W x H equirectangular preview of the driver's --preview output (main.c:297-312,
 * which uses W = 2R, H = R): per pixel render_sample_dir (render.c:228-251) in the
 * direction of its longitude/latitude centre, then render_pixel8; [H][W][4] when bg is
 * NULL, else [H][W][3].  Converted on the device; returns -1 for W or H < 1. */
int render_preview8(Render *r, const float *bg, int W, int H, unsigned char *out);

/* ---- discrete adjoint of fluid_step ----
 * Gradient of a loss L(omega_T) of the final vorticity after nsteps steps of fluid_step
 * with a prescribed dt schedule, with respect to the initial vorticity omega0 (taken
 * through set-up's mean removal, then a cold solve) and to the forcing strength
 * F->prm.forcing.  Needs an FP64-state fluid (GGB_PREC_FP64 or GGB_PREC_MIXED) and a
 * reciprocal stencil.  Checkpoints (omega, psi) every checkpoint_every steps; the backward
 * pass recomputes one segment at a time.  The Poisson solves (forward and adjoint) use
 * F->prm.cg_tol; the gradient is exact for exact solves.  Both calls overwrite the
 * fluid's state: after fluid_adjoint_forward it is the rollout's final state, after
 * fluid_adjoint_backward its initial state (omega0 after the mean removal, its psi, and
 * the velocity and maximum speed of that psi). */
typedef struct FluidAdjoint FluidAdjoint;
FluidAdjoint *fluid_adjoint_new(Fluid *F, int checkpoint_every);   /* NULL on error */
void fluid_adjoint_free(FluidAdjoint *A);
/* omega0 [n] in; omega_T [n] out (may be NULL); 0 on success */
int fluid_adjoint_forward(FluidAdjoint *A, const double *omega0, const double *dts, int nsteps,
			  double *omega_T);
/* g_omega_T = dL/d omega_T [n] in; g_omega0 [n] and *g_forcing out; 0 on success */
int fluid_adjoint_backward(FluidAdjoint *A, const double *g_omega_T, double *g_omega0, double *g_forcing);
/* CG iteration counts of the last fluid_adjoint_forward: its[0] of the set-up solve,
 * its[1 + s] of step s.  Returns the count (nsteps + 1), or -1 before any forward pass;
 * its may be NULL to query the count. */
int fluid_adjoint_iterations(const FluidAdjoint *A, int *its);
/* its != NULL: later forward passes run each of their count = nsteps + 1 Poisson solves
 * for exactly its[k] iterations instead of stopping at the tolerance, so the rollout is a
 * smooth map of omega0 for finite-difference checks (a tolerance test stops neighbouring
 * inputs after different counts: a jump of the solve's tolerance in the output).  NULL
 * restores the tolerance test.  0 on success, -1 for a negative count. */
int fluid_adjoint_fix_iterations(FluidAdjoint *A, const int *its, int count);

#ifdef __cplusplus
}
#endif

#endif
