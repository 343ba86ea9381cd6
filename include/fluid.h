/* Barotropic vorticity / stream-function solver on the cubed sphere, on one GPU.
 *
 *   d(omega)/dt + u . grad(omega) = forcing,   Laplacian(psi) = omega,   u = r x grad(psi)
 *
 * Drop-in for the reference's src/fluid.h:12-64: same entry points, same argument
 * meaning, same error behaviour (fluid_init returns 0; allocation failure inside a
 * step prints and exits).  The state lives on the GPU.  The float arrays omega,
 * omega_band, psi and vel are HOST MIRRORS of it:
 *   - fluid_set_state, fluid_solve_psi and fluid_compute_velocity behave like the
 *     reference on those arrays (they upload what they read and download what they
 *     write), so callers that poke F.omega / F.psi directly keep working;
 *   - fluid_step and fluid_max_speed work on the device state only; call
 *     fluid_sync_host() (ggb_ext.h) to refresh the mirrors after steps;
 *   - F.vel is also the handle particles_advect() recognises to read the device
 *     velocity without a copy.
 */
#ifndef GGB_FLUID_H
#define GGB_FLUID_H

#include "cubesphere.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Storage precision of the device state (extension; the zero default is FP64). */
enum {
	GGB_PREC_FP64 = 0,  /* every state array double; V-cycle double */
	GGB_PREC_MIXED = 1, /* state and CG double; multigrid V-cycle float */
	GGB_PREC_REF32 = 2, /* float storage with double per-cell arithmetic (the reference's) */
};

/* reference fluid.h:12-19, plus `precision` (appended, so positional initialisers of
 * the six reference fields still compile and select FP64). */
typedef struct {
	double forcing;  /* relaxation rate toward band vorticity (0 = off) */
	double rotation; /* planetary rotation rate for Coriolis (0 = off) */
	int bfecc;       /* BFECC advection with limiter */
	double cg_tol;   /* relative residual tolerance */
	int cg_maxit;
	int multigrid;   /* multigrid preconditioner (else Jacobi) */
	int precision;   /* GGB_PREC_* */
} FluidParams;

struct ggb_fluid_dev;

/* reference fluid.h:33-47 (the fields callers read; scratch lives on the device) */
typedef struct {
	const CubeSphere *cs;
	FluidParams prm;
	int nlev;           /* multigrid levels (fluid.c:44-46 rule) */
	float *omega;       /* host mirror: relative vorticity */
	float *omega_band;  /* host mirror: forcing target */
	float *psi;         /* host mirror: stream function */
	float *vel;         /* host mirror: [ncells*3] tangent velocity */
	int cg_iters;       /* iterations of the last solve */
	double cg_relres;   /* recurrence relative residual of the last solve */
	struct ggb_fluid_dev *dev;
} Fluid;

/* reference fluid.h:49-64 */
int fluid_init(Fluid *F, const CubeSphere *cs, const FluidParams *prm);
void fluid_free(Fluid *F);

void fluid_curl(const CubeSphere *cs, const float *vel, float *omega);
void fluid_divergence(const CubeSphere *cs, const float *vel, float *div);
void fluid_laplacian(const CubeSphere *cs, const float *x, float *y);
void fluid_remove_mean(const CubeSphere *cs, float *f);

void fluid_set_state(Fluid *F, const float *vel_initial, const float *vel_band);
int fluid_solve_psi(Fluid *F);
void fluid_compute_velocity(Fluid *F);
void fluid_step(Fluid *F, double dt);
double fluid_max_speed(const Fluid *F);

#ifdef __cplusplus
}
#endif

#endif
