"""The workloads other than the benchmarked one, at their own sizes:

  * N=64 plain method (24,576 cells, 100k particles, faces 256^2, FP64, no forcing or
    rotation, 4-octave noise, fade 0.02, 200 steps): the first 10 steps against the FP64
    oracle at the small cases' bars, then all 200 steps against the oracle (8-bit texels
    within 1 LSB on >= 99.9 %, fields and particles to a stated looser bound) and
    through size-independent properties (zero mean vorticity, no new extrema, unit
    particle radius, the solve's true residual), and a bitwise repeat of the GPU run;
  * N=256 drop-in (393,216 cells, 1M particles, faces 1024^2, FP64):
    the first 10 steps against the FP64 oracle on the same seeded inputs, to the same
    bars as the small parity cases (fields 1e-10 relative, particles 1e-9 rad, 8-bit
    texels within 1 LSB on >= 99.9 %, CG iteration counts equal);
  * N=128 adjoint (50 FP64 steps, checkpoints every 10, loss against the 50-step
    forward run of another noise seed): the gradient against the oracle's FP64 tape
    adjoint, and against central differences (h = 1e-6) on 64 random cells: their
    directional derivative to 1e-6, each cell to the FP64 noise floor of the rollout;
  * N=1024 (6,291,456 cells, 16M particles, faces 4096^2, mixed precision, forcing 0.2,
    rotation 5, 6-octave noise): a few steps through size-independent properties, and a
    short adjoint window on the same state.
"""
import numpy as np
import pytest

from conftest import levels_for
from test_gpu_parity import rel, run_more, run_pair

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def test_c1_drop_in_first_10_steps(ggb, orc, gpu):
    r = run_pair(ggb, orc, 256, 10, forcing=0.0, rotation=0.0, seed=1, octaves=4, cg_tol=1e-4,
                 particles=1_000_000, R=1024)
    F, O = r["F"], r["O"]
    assert r["its"] == r["oits"]
    for field in (ggb.FIELD_OMEGA, ggb.FIELD_PSI, ggb.FIELD_VEL):
        assert rel(F.get(field), O.get(field)) < 1e-10, field
    P, OP = r["P"], r["OP"]
    pos, prev = P.get()
    assert np.max(np.abs(pos - OP.pos)) < 1e-9
    assert np.max(np.abs(prev - OP.prev)) < 1e-9
    Rr, OR = r["R"], r["OR"]
    a, b = Rr.faces8().astype(int), OR.faces8().astype(int)
    assert np.max(np.abs(a - b)) <= 1
    assert np.count_nonzero(a != b) <= 0.001 * a[..., 0].size


def test_c0_plain_method_200_steps(ggb, orc, gpu):
    N, NP, R, steps = 64, 100_000, 256, 200
    r = run_pair(ggb, orc, N, 10, forcing=0.0, rotation=0.0, seed=1, octaves=4, cg_tol=1e-4,
                 particles=NP, R=R)
    F, O, P, OP, Rr, OR, g = r["F"], r["O"], r["P"], r["OP"], r["R"], r["OR"], r["g"]
    assert r["its"] == r["oits"]
    for field in (ggb.FIELD_OMEGA, ggb.FIELD_PSI, ggb.FIELD_VEL):
        assert rel(F.get(field), O.get(field)) < 1e-10, field
    pos, prev = P.get()
    assert np.max(np.abs(pos - OP.pos)) < 1e-9
    assert np.max(np.abs(prev - OP.prev)) < 1e-9
    a, b = Rr.faces8().astype(int), OR.faces8().astype(int)
    assert np.max(np.abs(a - b)) <= 1
    assert np.count_nonzero(a != b) <= 0.001 * a[..., 0].size

    area = g.area.astype(np.float64)
    om_start = O.get(0)
    lo, hi = om_start.min(), om_start.max()
    relres = []

    def check(o):
        om = o["F"].get(ggb.FIELD_OMEGA)
        # no forcing: zero mean, and the limited BFECC update makes no new extrema
        assert abs(np.sum(area * om)) <= 1e-12 * np.sum(area * np.abs(om))
        assert om.min() >= lo - 1e-9 and om.max() <= hi + 1e-9
        relres.append(o["F"].true_relres())

    r["on_step"] = check
    run_more(r, steps - 10)
    assert max(relres) <= 1.0001e-4, max(relres)  # the recurrence residual stops the solve
    its_equal = sum(int(a == b) for a, b in zip(r["its"], r["oits"]))
    errs = {f: rel(F.get(f), O.get(f)) for f in (ggb.FIELD_OMEGA, ggb.FIELD_PSI, ggb.FIELD_VEL)}
    pos, prev = P.get()
    perr = float(np.max(np.abs(pos - OP.pos)))
    a, b = Rr.faces8().astype(int), OR.faces8().astype(int)
    print("C0 after %d steps: CG counts equal on %d steps, fields %s, particles %.2e rad, texels "
          "differing %d of %d (max %d)" % (steps, its_equal, errs, perr, np.count_nonzero(a != b), a.size,
                                          np.max(np.abs(a - b))))
    assert np.max(np.abs(np.linalg.norm(pos.reshape(-1, 3), axis=1) - 1)) < 1e-12
    assert np.max(np.abs(a - b)) <= 1
    assert np.count_nonzero(a != b) <= 0.001 * a[..., 0].size
    for f, e in errs.items():
        assert e < 1e-6, (f, e)
    assert perr < 1e-6
    # the GPU run repeats bit for bit
    buf = Rr.sync_host().copy()
    om = F.get(ggb.FIELD_OMEGA)
    F2 = ggb.Fluid(g, forcing=0.0, rotation=0.0, cg_tol=1e-4)
    vel0, vband = g.initial_velocity(seed=1, octaves=4)
    F2.set_state(vel0, vband)
    P2 = ggb.Particles(NP)
    P2.color_from_image(ggb.strip_to_float(ggb.banded_strip(1)), 64, 1024)
    R2 = ggb.Render(R)
    assert len(r["dts"]) == steps
    for dt in r["dts"]:
        P2.advect(g, F2, dt)
        F2.step(dt)
        R2.fade(0.02)
        R2.deposit(P2)
    assert R2.g == Rr.g
    assert np.array_equal(F2.get(ggb.FIELD_OMEGA).view(np.uint64), om.view(np.uint64))
    assert np.array_equal(R2.sync_host().view(np.uint32), buf.view(np.uint32))


def _adjoint_case(ggb, orc, N=128, steps=50, seed_a=1, seed_b=2):
    """omega0 from noise seed A; target = the 50-step forward run from seed B."""
    levels = levels_for(ggb, orc, N)
    g = levels[0]
    O = orc.Oracle("f64")
    oms = []
    for seed in (seed_a, seed_b):
        vel0, vband = g.initial_velocity(seed=seed, octaves=3)
        oms.append(O.curl(g, vel0.astype(np.float64)))
    _, vband = g.initial_velocity(seed=seed_a, octaves=3)
    omb = O.curl(g, vband.astype(np.float64))
    omb -= np.sum(g.area * omb) / g.total_area
    return levels, g, oms[0], oms[1], omb


def test_c3_adjoint(ggb, orc, gpu):
    N, steps, K = 128, 50, 10
    levels, g, om0, om_b, omb = _adjoint_case(ggb, orc, N, steps)
    F = ggb.Fluid(g, forcing=0.5, rotation=10.0, cg_tol=1e-10, cg_maxit=2000, precision=ggb.PREC_FP64)
    F.set(ggb.FIELD_OMEGA_BAND, omb)
    A = ggb.FluidAdjoint(F, K)
    # dt frozen from the forward pass of omega0
    F.set(ggb.FIELD_OMEGA, om0)
    F.solve_psi_device()
    F.compute_velocity()
    dts = np.full(steps, g.h / max(F.max_speed(), 1e-6))
    target = A.forward(om_b, dts)
    area = g.area.astype(np.float64)

    def loss(w):
        return 0.5 * float(np.sum(area * (w - target) ** 2))

    wT = A.forward(om0, dts)
    g0, gk = A.backward(area * (wT - target))
    # the oracle's FP64 tape adjoint on the same inputs
    O = orc.OracleFluid("f64", levels, forcing=0.5, rotation=10.0, cg_tol=1e-10, cg_maxit=2000)
    O.set(1, omb)
    wTo = O.adj_forward(om0, dts)
    g0o, gko = O.adj_backward(area * (wTo - target))
    print("omega_T rel %.2e  gradient rel %.2e  dL/dk rel %.2e" %
          (rel(wT, wTo), rel(g0, g0o), abs(gk - gko) / abs(gko)))
    assert rel(wT, wTo) < 1e-8
    assert rel(g0, g0o) < 1e-6
    assert abs(gk - gko) <= 1e-6 * abs(gko)
    # Central differences (h = 1e-6) on 64 random cells.  The numerator L(w+) - L(w-) is
    # formed as 0.5 sum A (w+ - w-)(w+ + w- - 2 target), so no two large loss values are
    # subtracted.  What remains is the forward map's own FP64 rounding: the Poisson solve
    # spreads a one-cell perturbation over every cell, each rounded differently in the two
    # rollouts, which leaves about 1e-12 (1e-15 L) of noise in the numerator whatever h
    # is -- measured per-cell errors 6.5e-6 / 5.7e-5 / 6.0e-4 at h = 1e-5 / 1e-6 / 1e-7,
    # with every solve taking the same iteration count in all rollouts.  So:
    #   * the directional derivative over the 64 cells (d = sign(g) on them, 64 times one
    #     cell's signal) is checked at h = 1e-6 to 1e-6 relative;
    #   * each cell's derivative is checked to the noise floor, 1e-4 relative over the 64.
    rng = np.random.default_rng(7)
    cells = rng.choice(g.n, 64, replace=False)
    h = 1e-6

    def fd(d):
        wp = A.forward(om0 + h * d, dts)
        wm = A.forward(om0 - h * d, dts)
        return 0.5 * float(np.sum(area * (wp - wm) * (wp + wm - 2 * target))) / (2 * h)

    d = np.zeros(g.n)
    d[cells] = np.sign(g0[cells])
    fdd, ex = fd(d), float(g0 @ d)
    print("directional FD over 64 cells: %.12e vs adjoint %.12e, relative error %.2e" %
          (fdd, ex, abs(fdd - ex) / abs(ex)))
    assert abs(fdd - ex) <= 1e-6 * abs(ex)
    fds = []
    for c in cells:
        e = np.zeros(g.n)
        e[c] = 1.0
        fds.append(fd(e))
    fds = np.array(fds)
    r64 = rel(fds, g0[cells])
    print("per-cell FD on 64 cells: relative error %.2e (max abs %.2e, L = %.3e)" %
          (r64, np.max(np.abs(fds - g0[cells])), loss(wT)))
    assert r64 <= 1e-4
    A.free()
    F.free()


def test_c4_large_mixed(ggb, gpu):
    N, NP, R = 1024, 16_000_000, 4096
    g = ggb.Grid(N)
    vel0, vband = g.initial_velocity(seed=1, octaves=6)
    F = ggb.Fluid(g, forcing=0.2, rotation=5.0, cg_tol=1e-8, cg_maxit=500, precision=ggb.PREC_MIXED)
    F.set_state(vel0, vband)
    P = ggb.Particles(NP)
    P.color_from_image(ggb.strip_to_float(ggb.banded_strip(1)), 64, 1024)
    Rr = ggb.Render(R, 1.0, 0.35)
    area = g.area.astype(np.float64)
    for _ in range(3):
        dt = g.h / max(F.max_speed(), 1e-6)
        P.advect(g, F, dt)
        F.step(dt)
        Rr.fade(0.02)
        Rr.deposit(P)
        assert F.true_relres() <= 1e-8
        om = F.get(ggb.FIELD_OMEGA)
        assert abs(np.sum(area * om)) <= 1e-10 * np.sum(area * np.abs(om))
    pos, prev = P.get()
    r = np.linalg.norm(pos.reshape(-1, 3), axis=1)
    assert np.max(np.abs(r - 1)) < 1e-12
    buf = Rr.sync_host()
    assert np.all(np.isfinite(buf))
    assert Rr.last_entries() > NP
    # a 20-step adjoint window with checkpoints every 5, on the evolved state
    om0 = F.get(ggb.FIELD_OMEGA)
    dts = np.full(20, g.h / max(F.max_speed(), 1e-6))
    A = ggb.FluidAdjoint(F, 5)
    wT = A.forward(om0, dts)
    g0, gk = A.backward(area * (wT - om0))
    assert np.all(np.isfinite(g0)) and np.isfinite(gk)
    assert np.linalg.norm(g0) > 0
    A.free()
    F.free()
    P.free()
    Rr.free()
