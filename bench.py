"""Benchmark of the per-step hot path on one MI355X.

Workload (default): N=512 cubed sphere (1,572,864 cells), 4,000,000 particles, 2048^2
faces, forcing 0.5, rotation 10, CG tolerance 1e-8, mixed precision (FP32 V-cycle in
FP64 CG), 8 bands + 4-octave curl noise from seed 1, cfl 1, fade 0.02, trail alpha 0.35,
splat radius 1.  A step is the driver loop of the reference (main.c:421-435): max speed
-> dt, particles_advect, fluid_step, render_fade + render_deposit.  By default the timed
window is the workload's whole run, 2000 steps after 5 warm-up steps: the flow develops
zonal jets over the first few hundred steps, trails lengthen and overlap, and the
render's cost per step roughly doubles, so a short early window would not represent it.

Reports (one JSON line): fluid steps/s (mixed; the headline `value`), Msplat-particles/s,
the FP64 path's fluid steps/s on the same workload, the true FP64 relative residual of
both, the roofline of the dominant kernel, the draw (HIP events on the library stream over
the timed steps), and the CPU baseline timed on this host.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gaseous-gusty-ball_amd"))

import ggb  # noqa: E402

CONFIGS = {
    # name: N, particles, R, forcing, rotation, cg_tol, precision, octaves
    "c2": dict(N=512, P=4_000_000, R=2048, forcing=0.5, rotation=10.0, cg_tol=1e-8,
               precision=ggb.PREC_MIXED, octaves=4),
    "c1": dict(N=256, P=1_000_000, R=1024, forcing=0.0, rotation=0.0, cg_tol=1e-4,
               precision=ggb.PREC_FP64, octaves=4),
    "c0": dict(N=64, P=100_000, R=256, forcing=0.0, rotation=0.0, cg_tol=1e-4,
               precision=ggb.PREC_FP64, octaves=4),
    "c4": dict(N=1024, P=16_000_000, R=4096, forcing=0.2, rotation=5.0, cg_tol=1e-8,
               precision=ggb.PREC_MIXED, octaves=6),
}
HBM_PEAK = 8.0e12
FP64_PEAK = 78.6e12  # MI355X FP64 vector (non-matrix)

# Probed kernels (ids of the library's launch probes) and their algorithmic bytes per
# launch under the per-pass model: every pass reads its inputs and writes its outputs
# once, gathered neighbours are cache hits, the Laplacian stencil is one 8-byte record
# per cell (its class weights are cache-resident), transfer coefficients are the stored
# float tables.  n = cells, P = particles, E = (particle, tile) entries, T = texels
# touched, Ts/Tm = state / multigrid word.
PROBES = {
    "k_render_draw": 1, "render_count": 2, "mg_up(level 0)": 3, "k_mg_down(level 0)": 4,
    "k_cg_Ap": 5, "k_cg_update_down": 6, "k_particles_advect": 7, "k_departures": 8,
    "k_render_fill": 9,
}


def kernel_bytes(name, n, P, E, T, Ts, Tm, Tz, reorder_frac=1.0, cheb=False):
    if name == "k_render_draw":      # the render contract: prev, pos, colour per particle (f32) + texel RMW
        return 40 * P + 32 * T
    if name == "render_count":       # current position (FP64) for the home tile
        return P * 24
    if name == "k_render_fill":      # endpoints read, sample count, list; on reordering deposits
        return P * (48 + 2) + reorder_frac * P * (20 + 68) + E * 4  # also colour+index, copy out
    if name == "mg_up(level 0)":
        if cheb:                     # r, x_pre, coarse x, stencil record, dinv, Ap in; z, A z out
            return n * (Ts + 4 + 1 + 8 + 4 + Ts + 4 + Ts)
        return n * (Ts + 4 + 8 + 16 + Tm / 4 + Tm)   # b, dinv, stencil record, tap weights, coarse x, z out
    if name == "k_mg_down(level 0)":
        if cheb:                     # r, stencil record, dinv; x_pre and the coarse b out
            return n * (Ts + 8 + 4 + 4 + 1)
        return n * (Ts + 4 + 8 + Tm)  # b, dinv, stencil record, r1 out
    if name == "k_cg_Ap":            # z, p in, stencil record, p out, Ap out
        return n * (Tz + Ts + 8 + 2 * Ts)
    if name == "k_cg_update_down":
        if cheb:                     # r, A z, Ap, p, psi, z, stencil record, dinv in; psi, p, Ap, r, x_pre, coarse b out
            return n * (5 * Ts + 4 + 8 + 4 + 4 * Ts + 4 + 1)
        return n * (6 * Ts + 12 + Tm)  # psi, r, p, Ap, dinv, stencil record in; psi, r, r1 out
    if name == "k_cg_update_cheb":   # the update alone: r, A z, Ap, p, psi, z in; psi, p, Ap, r out
        return n * (5 * Ts + 4 + 4 * Ts)
    if name == "k_particles_advect":  # pos in, pos+prev out (FP64), velocity field once
        return P * 72 + n * 3 * Ts
    if name == "k_departures":       # vel, omega, centre in; q0, dep_b, dep_f out
        return n * (3 * Ts + Ts + 12 + Ts + 6 * Ts)
    raise KeyError(name)


def fluid_step_bytes(N, its, vcycles, precision, cheb, galerkin=True):
    """Algorithmic bytes of one fluid step under the per-pass model (each pass reads its
    inputs and writes its outputs once; gathered neighbours, the stencil's class weights
    and the small transfer tables are cache hits): the fixed passes plus, per V-cycle, the
    level-0 down and up passes and every coarser level's, and per CG iteration the Ap pass
    (FP64 mode).  The mixed PCG runs its iterations' updates inside the V-cycles' first
    passes (one V-cycle per iteration, the last update alone); its coarse levels read their
    Galerkin operators' 24 half-precision weights per cell.  Returns (bytes, breakdown)."""
    n = 6 * N * N
    Ts = 4 if precision == ggb.PREC_REF32 else 8
    Tm = 8 if precision == ggb.PREC_FP64 else 4
    fixed = {
        "departures": n * 100 if Ts == 8 else n * 52,
        "advect + bfecc + limiter": n * (40 + 48 + 64) * Ts // 8,
        "cg_start": n * (3 * Ts + 4 + 8),
        "velocity + mean": n * (2 * Ts + 96 + 32 + 12 + 3 * Ts + 12),
        "extrapolated start": n * 5 * Ts if precision == ggb.PREC_MIXED else 0,
    }
    per_cycle = 0.0
    levels = []
    m = n
    Nl = N
    while Nl >= 4 and m > 1536:
        levels.append(m)
        if Nl % 2:
            break
        Nl //= 2
        m = 6 * Nl * Nl
    for l, nl in enumerate(levels):
        if l == 0:
            continue
        if cheb and galerkin:
            # a Galerkin level's down pass (b, 48 B of weights, dinv in; x_pre and the next
            # level's b out) and up pass (b, x_pre, coarse x, weights, dinv in; x out)
            per_cycle += nl * 126
        elif cheb:
            # a coarse level's down pass (b, stencil record, dinv in; x_pre and the next level's
            # b out) and up pass (b, x_pre, coarse x, stencil record, dinv in; x out)
            per_cycle += nl * 46
        else:
            # the same with the restriction and prolongation tables read
            per_cycle += nl * (4.5 * Tm + 56)
    if m <= 1536:
        per_cycle += Tm * m * m + 2 * Tm * m  # the dense bottom operator and its vectors
    up0 = kernel_bytes("mg_up(level 0)", n, 0, 0, 0, Ts, Tm, Tm, cheb=cheb)
    down0 = kernel_bytes("k_mg_down(level 0)", n, 0, 0, 0, Ts, Tm, Tm, cheb=cheb)
    upd0 = kernel_bytes("k_cg_update_down", n, 0, 0, 0, Ts, Tm, Tm, cheb=cheb)
    if cheb:
        # its V-cycles (the first without an update), its - 1 fused updates, one alone
        alone = kernel_bytes("k_cg_update_cheb", n, 0, 0, 0, Ts, Tm, Tm)
        solver = {"V-cycles": its * (per_cycle + up0) + down0 + (its - 1) * upd0 + alone}
    else:
        ap = kernel_bytes("k_cg_Ap", n, 0, 0, 0, Ts, Tm, Tm)
        solver = {
            "V-cycles": vcycles * (per_cycle + up0) + down0 + (vcycles - 1) * upd0,
            "Ap": its * ap,
        }
    b = dict(fixed)
    b.update(solver)
    return float(sum(b.values())), {k: int(v) for k, v in b.items()}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(cfg, precision, with_particles=True, seed=1):
    g = ggb.Grid(cfg["N"])
    vel0, vband = g.initial_velocity(seed=seed, octaves=cfg["octaves"])
    F = ggb.Fluid(g, forcing=cfg["forcing"], rotation=cfg["rotation"], cg_tol=cfg["cg_tol"],
                  cg_maxit=500, precision=precision)
    F.set_state(vel0, vband if cfg["forcing"] > 0 else None)
    P = R = None
    if with_particles:
        P = ggb.Particles(cfg["P"])
        P.color_from_image(ggb.strip_to_float(ggb.banded_strip(seed)), 64, 1024)
        R = ggb.Render(cfg["R"], 1.0, 0.35)
    return g, F, P, R


class PhaseTimes:
    """HIP events on the library stream around each phase of each step.  The elapsed times
    are read only after the run, so timing adds no host waits (a per-phase event
    synchronisation would drain the queue between phases, which a driver does not)."""

    def __init__(self):
        self.pool, self.pairs, self.open = [], {k: [] for k in "pfr"}, None

    def _event(self):
        return self.pool.pop() if self.pool else ggb.lib().ggbk_event_new()

    def start(self):
        self.open = self._event()
        ggb.lib().ggbk_event_record(self.open)

    def stop(self, phase):
        b = self._event()
        ggb.lib().ggbk_event_record(b)
        self.pairs[phase].append((self.open, b))

    def collect(self, acc):
        """Adds each phase's total milliseconds to acc and recycles the events."""
        L = ggb.lib()
        for k, pairs in self.pairs.items():
            for a, b in pairs:
                acc[k] += L.ggbk_event_ms(a, b)
                self.pool += [a, b]
            pairs.clear()


def run_steps(g, F, P, R, nsteps, timers=None, acc=None):
    """Driver loop (reference main.c:427-451); per-phase device time from HIP events."""
    timers = timers or PhaseTimes()
    acc = acc if acc is not None else dict(p=0.0, f=0.0, r=0.0, its=0, entries=0)
    for _ in range(nsteps):
        vmax = F.max_speed()
        dt = 1.0 * g.h / max(vmax, 1e-6)
        if P is not None:
            timers.start()
            P.advect(g, F, dt)
            timers.stop("p")
        timers.start()
        F.step(dt)
        timers.stop("f")
        acc["its"] += F.cg_iters
        if R is not None:
            timers.start()
            R.fade(0.02)
            R.deposit(P)
            timers.stop("r")
            acc["entries"] += R.last_entries()
    timers.collect(acc)


def cpu_baseline(cfg, steps):
    """The FP64 restatement (fluid) and the reference's own optimised build (particles and
    render), on this host's cores, over a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ggb_oracle as O

    # every core this process may run on, as many as OMP_NUM_THREADS allows (a one-GPU
    # share of the host's cores where the host sets it)
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = min(avail, omp) if omp > 0 else avail
    O.set_omp_threads(cores)
    N = cfg["N"]
    nl = O.olib("f64").orc_mg_nlev(N)
    levels = [ggb.Grid(N >> l) for l in range(nl)]
    g = levels[0]
    vel0, vband = g.initial_velocity(seed=1, octaves=cfg["octaves"])
    OF = O.OracleFluid("f64", levels, forcing=cfg["forcing"], rotation=cfg["rotation"],
                       cg_tol=cfg["cg_tol"], cg_maxit=500)
    OF.set_state(vel0, vband if cfg["forcing"] > 0 else None)
    res = {"cores": cores, "fluid_s": 0.0, "render_s": 0.0, "particles_s": 0.0}
    have_ref = O.available("ref_fast")
    if have_ref:
        rg = O.RefGrid(N, kind="ref_fast")
        RP = O.RefParticles(cfg["P"], kind="ref_fast")
        RP.color_from_image(ggb.strip_to_float(ggb.banded_strip(1)), 64, 1024)
        RR = O.RefRender(cfg["R"], 1.0, 0.35, kind="ref_fast")
    for _ in range(steps):
        dt = g.h / max(OF.max_speed(), 1e-6)
        if have_ref:
            vel32 = OF.get(3).astype(np.float32)
            t0 = time.perf_counter()
            RP.advect(rg, vel32, dt)
            res["particles_s"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        OF.step(dt)
        res["fluid_s"] += time.perf_counter() - t0
        if have_ref:
            t0 = time.perf_counter()
            RR.fade(0.02)
            RR.deposit(RP)
            res["render_s"] += time.perf_counter() - t0
    res["steps"] = steps
    res["have_ref"] = have_ref
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    res["cpu_model"] = model
    res["host_cpus"] = os.cpu_count()
    res["affinity_cpus"] = avail
    return res


def reference_c1(steps=10):
    """The reference's own fluid (its -O3 build: float PCG, tol 1e-4) at N=256 on this
    host's cores, from the same seeded initial flow: fluid steps/s over `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ggb_oracle as O
    if not O.available("ref_fast"):
        return None
    cfg = CONFIGS["c1"]
    g = ggb.Grid(cfg["N"])
    vel0, _ = g.initial_velocity(seed=1, octaves=cfg["octaves"])
    rg = O.RefGrid(cfg["N"], kind="ref_fast")
    RF = O.RefFluid(rg, forcing=0.0, rotation=0.0, cg_tol=cfg["cg_tol"], cg_maxit=500)
    RF.set_state(vel0)
    t = 0.0
    its = 0
    for _ in range(steps):
        dt = g.h / max(RF.max_speed(), 1e-6)
        t0 = time.perf_counter()
        RF.step(dt)
        t += time.perf_counter() - t0
        its += RF.cg_iters
    return {"fluid_steps_per_s": round(steps / t, 3), "cg_iters_per_step": its / steps, "steps": steps}


def gpu_c1(steps=200, warmup=5):
    """The FP64 drop-in on the reference's default grid (N=256, tol 1e-4): fluid steps/s."""
    cfg = CONFIGS["c1"]
    g, F, _, _ = setup(cfg, ggb.PREC_FP64, with_particles=False)
    timers = PhaseTimes()
    acc = dict(p=0.0, f=0.0, r=0.0, its=0, entries=0)
    run_steps(g, F, None, None, warmup, timers, dict(acc))
    run_steps(g, F, None, None, steps, timers, acc)
    F.free()
    return {"fluid_steps_per_s": round(1000.0 / (acc["f"] / steps), 2), "cg_iters_per_step": acc["its"] / steps,
            "steps": steps}


def hbm_used_bytes():
    """Device memory in use (total - free, hipMemGetInfo), or None."""
    try:
        hip = C.CDLL("libamdhip64.so")
        free, total = C.c_size_t(), C.c_size_t()
        if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
            return None
        return total.value - free.value
    except OSError:
        return None


def precision_leg(cfg, checkpoints=(10, 100, 2000), seed=1):
    """Mixed (FP32 V-cycle in FP64 CG) against FP64 on the same workload: the fluids and
    their particles stepped in lockstep with the FP64 run's dt; relative L2 errors of
    omega, psi and u and the largest particle position error (rad) at the checkpoints.
    A control run, FP64 at a ten times tighter tolerance, gives the same differences for
    FP64 itself: the flow amplifies any change within the solve's tolerance, so after a
    few hundred steps the two FP64 runs part as far as mixed and FP64 do."""
    g = ggb.Grid(cfg["N"])
    vel0, vband = g.initial_velocity(seed=seed, octaves=cfg["octaves"])
    runs = []
    for prec, tol in ((ggb.PREC_FP64, cfg["cg_tol"]), (ggb.PREC_MIXED, cfg["cg_tol"]),
                      (ggb.PREC_FP64, 0.1 * cfg["cg_tol"])):
        F = ggb.Fluid(g, forcing=cfg["forcing"], rotation=cfg["rotation"], cg_tol=tol,
                      cg_maxit=500, precision=prec)
        F.set_state(vel0, vband if cfg["forcing"] > 0 else None)
        runs.append((F, ggb.Particles(cfg["P"])))
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out = {}
    for step in range(1, max(checkpoints) + 1):
        dt = g.h / max(runs[0][0].max_speed(), 1e-6)
        for F, P in runs:
            P.advect(g, F, dt)
            F.step(dt)
        if step in checkpoints:
            (F64, P64) = runs[0]
            p64 = P64.get()[0]
            for key, (Fo, Po) in (("mixed", runs[1]), ("fp64_tol_tenth", runs[2])):
                e = {}
                for name, fld in (("omega", ggb.FIELD_OMEGA), ("psi", ggb.FIELD_PSI), ("u", ggb.FIELD_VEL)):
                    e[name] = rel(Fo.get(fld), F64.get(fld))
                e["particles_max_rad"] = float(np.max(np.abs(Po.get()[0] - p64)))
                out.setdefault(key, {})[str(step)] = e
            log("at step %d: mixed vs fp64 %s; fp64 at tol/10 vs fp64 %s" %
                (step, out["mixed"][str(step)], out["fp64_tol_tenth"][str(step)]))
    for F, P in runs:
        P.free()
        F.free()
    return out


def adjoint_window(cfg, steps=100, K=25, seed=1):
    """Adjoint window of the large workload: forward and backward over `steps` steps of
    the mixed-precision fluid with checkpoints every K, timed."""
    g = ggb.Grid(cfg["N"])
    vel0, vband = g.initial_velocity(seed=seed, octaves=cfg["octaves"])
    F = ggb.Fluid(g, forcing=cfg["forcing"], rotation=cfg["rotation"], cg_tol=1e-10, cg_maxit=2000,
                  precision=cfg["precision"])
    F.set_state(vel0, vband)
    om0 = F.get(ggb.FIELD_OMEGA)
    dts = np.full(steps, g.h / max(F.max_speed(), 1e-6))
    A = ggb.FluidAdjoint(F, K)
    area = g.area.astype(np.float64)
    t0 = time.perf_counter()
    wT = A.forward(om0, dts)
    t1 = time.perf_counter()
    g0, gk = A.backward(area * (wT - om0))
    t2 = time.perf_counter()
    res = {"workload": "N=%d, %d steps, checkpoint every %d, CG tol 1e-10" % (cfg["N"], steps, K),
           "forward_ms": round((t1 - t0) * 1000, 1), "backward_ms": round((t2 - t1) * 1000, 1),
           "grad_finite": bool(np.all(np.isfinite(g0))), "dL_dforcing": gk}
    A.free()
    F.free()
    return res


def adjoint_leg(N=128, steps=50, K=10, seed=1, fd_steps=10):
    """Discrete adjoint at N=128 over 50 FP64 steps with checkpoints every 10: timed
    forward and backward; and a central-difference check of the gradient along one random
    direction over the first 10 steps (best of steps 1e-5, 1e-6, 1e-7).  Over longer
    horizons the forward map's kinks (the BFECC limiter's clamps, bilinear taps changing
    cell) make central differences themselves inaccurate: 2e-3 at 50 steps and step 1e-6,
    1.3e-7 at 10 steps."""
    g = ggb.Grid(N)
    vel0, vband = g.initial_velocity(seed=seed, octaves=3)
    om0 = g.curl_f64(vel0.astype(np.float64))
    omb = g.curl_f64(vband.astype(np.float64))
    omb -= np.sum(g.area * omb) / g.total_area
    F = ggb.Fluid(g, forcing=0.5, rotation=10.0, cg_tol=1e-10, cg_maxit=2000, precision=ggb.PREC_FP64)
    F.set(ggb.FIELD_OMEGA_BAND, omb)
    F.set_state(vel0, vband)
    dts = np.full(steps, g.h / max(F.max_speed(), 1e-6))
    rng = np.random.default_rng(seed)
    target = 0.5 * om0[rng.permutation(g.n)]
    A = ggb.FluidAdjoint(F, K)
    area = g.area.astype(np.float64)

    def loss(w):
        return 0.5 * float(np.sum(area * (w - target) ** 2))

    A.forward(om0, dts)                     # warm-up (allocations, first solves)
    t0 = time.perf_counter()
    wT = A.forward(om0, dts)
    t1 = time.perf_counter()
    g0, gk = A.backward(area * (wT - target))
    t2 = time.perf_counter()
    d = rng.standard_normal(g.n)
    fdts = dts[:fd_steps]
    wT = A.forward(om0, fdts)
    g0s, _ = A.backward(area * (wT - target))
    ad = float(g0s @ d)
    errs = []
    for eps in (1e-5, 1e-6, 1e-7):
        fd = (loss(A.forward(om0 + eps * d, fdts)) - loss(A.forward(om0 - eps * d, fdts))) / (2 * eps)
        errs.append(abs(fd - ad) / abs(ad))
    A.free()
    F.free()
    return {"workload": "N=%d (%d cells), %d FP64 steps, checkpoint every %d, CG tol 1e-10" % (N, g.n, steps, K),
            "forward_ms": round((t1 - t0) * 1000, 2), "backward_ms": round((t2 - t1) * 1000, 2),
            "backward_over_forward": round((t2 - t1) / (t1 - t0), 2), "dL_dforcing": gk,
            "fd_check": "first %d steps, one random direction, best of steps 1e-5/1e-6/1e-7" % fd_steps,
            "fd_directional_rel_err": min(errs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the default timed window is the workload's whole run (2000 steps at N=512): the
    # flow develops jets over the first hundreds of steps, and the render's cost with them
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the N=256 drop-in leg")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the mixed-vs-FP64 error run (steps 10, 100, 2000)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain copy kernel of MIB MiB (counter calibration)")
    # the draw has the most device time over a whole run (a quarter of it at N=512);
    # over the warm-up's few early steps a multigrid kernel can come out ahead, so the
    # warm-up's choice ("auto") is not the default
    ap.add_argument("--roofline-kernel", default="k_render_draw",
                    choices=sorted(PROBES) + ["auto"],
                    help="kernel whose roofline is reported (auto: most time in the warm-up)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(the path would shard by cube face, exchanging one ring of seam cells per stencil "
              "pass and two scalars per CG iteration)", file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    if ggb.device_init() != 0:
        sys.exit("bench.py: no GPU")
    if args.calib_copy:
        # a plain copy of known size, so counter runs can scale FETCH_SIZE / WRITE_SIZE
        nb = args.calib_copy << 20
        a, b = ggb.lib().ggbk_malloc(nb), ggb.lib().ggbk_malloc(nb)
        ggb.lib().ggbk_copy_kernel(b, a, nb)
        ggb.lib().ggb_sync()
        ggb.lib().ggbk_free(a)
        ggb.lib().ggbk_free(b)
        log("calibration copy: %d MiB read, %d MiB written" % (args.calib_copy, args.calib_copy))

    t0 = time.perf_counter()
    g, F, P, R = setup(cfg, cfg["precision"])
    log("setup %.1fs: N=%d cells=%d particles=%d R=%d prec=%d cold CG %d its" %
        (time.perf_counter() - t0, cfg["N"], g.n, cfg["P"], cfg["R"], cfg["precision"], F.cg_iters))
    timers = PhaseTimes()
    acc = dict(p=0.0, f=0.0, r=0.0, its=0, entries=0)
    L = ggb.lib()
    # warmup: one plain step (first launches load code objects), then the rest with every
    # probe on: per-kernel device time per step, to pick the dominant kernel
    nplain = 1 if args.warmup > 1 else 0
    run_steps(g, F, P, R, nplain, timers, dict(acc))
    for pid in PROBES.values():
        L.ggbk_probe_enable(pid, 1)
    run_steps(g, F, P, R, args.warmup - nplain, timers, dict(acc))
    warm = {}
    for name, pid in PROBES.items():
        ms = C.c_double()
        cnt = L.ggbk_probe_collect(pid, C.byref(ms))
        warm[name] = (cnt, ms.value)
        L.ggbk_probe_enable(pid, 0)
    top = max(warm, key=lambda k: warm[k][1])
    dominant = top if args.roofline_kernel == "auto" else args.roofline_kernel
    log("warmup done (%d steps); most warm-up time %s; roofline kernel %s" % (args.warmup, top, dominant))
    # timed region: only the dominant kernel is bracketed by events
    L.ggbk_probe_enable(PROBES[dominant], 1)
    t0 = time.perf_counter()
    run_steps(g, F, P, R, args.steps, timers, acc)
    ggb.lib().ggb_sync()
    wall = time.perf_counter() - t0
    dms = C.c_double()
    dcount = L.ggbk_probe_collect(PROBES[dominant], C.byref(dms))
    L.ggbk_probe_enable(PROBES[dominant], 0)
    hbm_peak = hbm_used_bytes()
    rstats = R.last_stats()
    K = args.steps
    ms_fluid = acc["f"] / K
    ms_render = acc["r"] / K
    ms_part = acc["p"] / K
    fluid_sps = 1000.0 / ms_fluid
    msplat = cfg["P"] / (ms_render / 1000.0) / 1e6
    relres = F.true_relres()
    log("timed %d steps in %.2fs: fluid %.3f ms (%.1f its/step), particles %.3f ms, render %.3f ms"
        % (K, wall, ms_fluid, acc["its"] / K, ms_part, ms_render))

    # roofline of the dominant kernel
    Ts = 4 if cfg["precision"] == ggb.PREC_REF32 else 8
    Tm = 8 if cfg["precision"] == ggb.PREC_FP64 else 4
    E = acc["entries"] / K  # (particle, tile) entries: mean over the timed steps
    T = rstats[8]           # texels touched: the last deposit's
    nbytes = kernel_bytes(dominant, g.n, cfg["P"], E, T, Ts, Tm, Tm)
    # the draw runs once per deposit, or twice when a list overflowed its bucket (the first
    # launch then returns at once): its time is taken per deposit
    per = K if dominant == "k_render_draw" else max(dcount, 1)
    avg_s = dms.value / 1000.0 / per
    achieved = nbytes / avg_s / 1e9
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    # recorded in profiles/traffic.json by the counter runs of this workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    tjson = json.load(open(tpath)).get(args.config, {}) if os.path.exists(tpath) else {}
    if tjson.get(dominant):
        traffic = tjson[dominant].get("bytes_per_launch")
    cheb = cfg["precision"] == ggb.PREC_MIXED and os.environ.get("GGB_MIXED_CHEB") != "0" and cfg["N"] % 16 == 0
    its = acc["its"] / K
    galerkin = cheb and os.environ.get("GGB_MIXED_GALERKIN") != "0"
    fbytes, fparts = fluid_step_bytes(cfg["N"], its, its + 1, cfg["precision"], cheb, galerkin)
    fach = fbytes / (ms_fluid / 1000.0)
    W = 8 if cfg["precision"] != ggb.PREC_REF32 else 4
    survey = (47 * W + 44 + 16 * 6 + (13 * W + 24 + (4.0 / 3.0) * (14.5 * (4 if cfg["precision"] != ggb.PREC_FP64
                                                                          else 8) + 40 + 48)) * its) * g.n
    kernels = {}
    # share of the probed deposits that reordered the particles (those run the count)
    rfrac = warm.get("render_count", (0, 0))[0] / max(warm.get("k_render_fill", (0, 0))[0], 1)
    for name, (cnt, ms) in warm.items():
        if cnt:
            b = kernel_bytes(name, g.n, cfg["P"], E, T, Ts, Tm, Tm, rfrac, cheb=cheb)
            avg = ms / 1000.0 / cnt
            kernels[name] = {"launches_per_step": round(cnt / max(args.warmup - nplain, 1), 2),
                             "avg_us": round(avg * 1e6, 2),
                             "algo_bytes": int(b),
                             "achieved_GBps": round(b / avg / 1e9, 1)}
            tk = tjson.get(name, {})
            if tk.get("bytes_per_launch"):
                kernels[name]["traffic_bytes"] = tk["bytes_per_launch"]
                kernels[name]["traffic_GBps"] = round(tk["bytes_per_launch"] / avg / 1e9, 1)
            if tk.get("fp64_flops_per_launch"):
                # FP64 VALU rate: counted flops per launch (committed counter profile) over
                # the live average duration, against the MI355X FP64 vector peak
                tf = tk["fp64_flops_per_launch"] / avg / 1e12
                kernels[name]["fp64_valu"] = {"flops_per_launch": tk["fp64_flops_per_launch"],
                                              "achieved": round(tf, 2), "peak": FP64_PEAK / 1e12,
                                              "unit": "TFLOP/s", "frac": round(tf * 1e12 / FP64_PEAK, 4)}
    out = {
        "metric": "fluid steps/s @N=512 + Msplat-particles/s; mixed vs FP64 at 1e-8 residual",
        "value": round(fluid_sps, 3),
        "unit": "fluid steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if cfg["precision"] == ggb.PREC_MIXED else "f64",
        "data": "synthetic (seeded curl noise + zonal bands; Fibonacci particles; seeded banded strip)",
        "config": {
            "workload": "cubed-sphere N=%d (%d cells), %d particles, faces %d^2, forcing %g, rotation %g, "
                        "CG tol %g, %s" % (cfg["N"], g.n, cfg["P"], cfg["R"], cfg["forcing"], cfg["rotation"],
                                           cfg["cg_tol"], "FP32 V-cycle in FP64 CG" if cfg["precision"] == 1
                                           else "FP64"),
            "parallelism": "1 GPU",
        },
        "msplat_particles_per_s": round(msplat, 2),
        "mparticles_advect_per_s": round(cfg["P"] / (ms_part / 1000.0) / 1e6, 2),
        "full_loop_steps_per_s": round(K / wall, 3),
        "cg_iters_per_step": round(acc["its"] / K, 2),
        "true_relres": relres,
        "splat_entries_per_step": acc["entries"] // K,
        "texels_touched_last_step": int(T),
        "splat_longest_tile_list": int(rstats[1]),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                     "traffic": traffic, "launches": dcount, "avg_us": round(avg_s * 1e6, 2),
                     "algo_bytes_per_launch": int(nbytes)},
        "fluid_roofline": {"bound": "hbm", "scope": "every pass of the fluid step", "achieved": round(fach / 1e9, 1),
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(fach / HBM_PEAK, 4),
                           "algo_bytes_per_step": int(fbytes), "passes": fparts,
                           "survey_model_bytes_per_step": int(survey),
                           "survey_model_frac": round(survey / (ms_fluid / 1000.0) / HBM_PEAK, 4)},
        "kernels_warmup": kernels,
    }
    if hbm_peak is not None:
        out["hbm_used_GB"] = round(hbm_peak / 1e9, 2)
    F.free()
    P.free()
    R.free()

    if not args.no_fp64 and cfg["precision"] == ggb.PREC_MIXED:
        g2, F2, _, _ = setup(cfg, ggb.PREC_FP64, with_particles=False)
        acc2 = dict(p=0.0, f=0.0, r=0.0, its=0, entries=0)
        run_steps(g2, F2, None, None, args.warmup, timers, dict(acc2))
        run_steps(g2, F2, None, None, K, timers, acc2)
        fp64_sps = 1000.0 / (acc2["f"] / K)
        out["fp64_fluid_steps_per_s"] = round(fp64_sps, 3)
        out["fp64_cg_iters_per_step"] = round(acc2["its"] / K, 2)
        out["fp64_true_relres"] = F2.true_relres()
        out["mixed_over_fp64"] = round(fluid_sps / fp64_sps, 3)
        F2.free()

    if not args.no_precision and cfg["precision"] == ggb.PREC_MIXED and args.config == "c2":
        t0 = time.perf_counter()
        pl = precision_leg(cfg)
        out["mixed_vs_fp64_error"] = pl["mixed"]
        out["fp64_tol_tenth_vs_fp64_error"] = pl["fp64_tol_tenth"]
        log("precision leg %.1fs" % (time.perf_counter() - t0))

    if not args.no_adjoint:
        t0 = time.perf_counter()
        if args.config == "c4":
            out["adjoint_window"] = adjoint_window(cfg)
        else:
            out["adjoint"] = adjoint_leg()
        log("adjoint leg %.1fs" % (time.perf_counter() - t0))

    if not args.no_c1:
        # the reference's own default grid and tolerance: the FP64 drop-in on the GPU, and
        # (with the CPU leg) the reference's own float fluid on this host
        t0 = time.perf_counter()
        out["c1_drop_in"] = {"workload": "N=256 (393216 cells), CG tol 1e-4, no forcing or rotation",
                             "gpu_fp64": gpu_c1()}
        if not args.no_cpu:
            out["c1_drop_in"]["reference_cpu"] = reference_c1()
        log("c1 leg %.1fs" % (time.perf_counter() - t0))

    if not args.no_cpu:
        t0 = time.perf_counter()
        cb = cpu_baseline(cfg, args.cpu_steps)
        cpu_fluid = cb["steps"] / cb["fluid_s"]
        out["cpu_baseline"] = {
            "value": round(cpu_fluid, 4),
            "unit": "fluid steps/s",
            "cores": cb["cores"],
            "kind": "port",
            "sample": "%d steps of the same workload after setup: fluid by the FP64 C restatement of the "
                      "reference (the reference's float PCG cannot reach tol 1e-8); particles and render by "
                      "the reference's own -O3 build; OpenMP on %d threads (the cores this process may use, "
                      "capped by OMP_NUM_THREADS; the host shows %d) of %s" %
                      (cb["steps"], cb["cores"], cb["host_cpus"], cb["cpu_model"]),
        }
        if cb["have_ref"]:
            out["cpu_baseline"]["msplat_particles_per_s"] = round(cfg["P"] * cb["steps"] / cb["render_s"] / 1e6, 3)
            out["cpu_baseline"]["mparticles_advect_per_s"] = round(cfg["P"] * cb["steps"] / cb["particles_s"] / 1e6, 3)
        log("cpu baseline %.1fs" % (time.perf_counter() - t0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
