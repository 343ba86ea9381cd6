"""The workloads other than the benchmarked one, at their own sizes:

  * N=256 drop-in (393,216 cells, 1M particles, faces 1024^2, FP64):
    the first 10 steps against the FP64 oracle on the same seeded inputs, to the same
    bars as the small parity cases (fields 1e-10 relative, particles 1e-9 rad, 8-bit
    texels within 1 LSB on >= 99.9 %, CG iteration counts equal);
  * N=128 adjoint (50 FP64 steps, checkpoints every 10, loss against the 50-step
    forward run of another noise seed): the gradient against the oracle's FP64 tape
    adjoint, and against central differences (h = 1e-6) on 64 random cells;
  * N=1024 (6,291,456 cells, 16M particles, faces 4096^2, mixed precision, forcing 0.2,
    rotation 5, 6-octave noise): a few steps through size-independent properties, and a
    short adjoint window on the same state.
"""
import numpy as np
import pytest

from conftest import levels_for
from test_gpu_parity import rel, run_pair

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
    # central differences (h = 1e-6) on 64 random cells.  Their floor is the rounding of
    # the loss itself, about 1e-16 L / h: measured 2e-7 median against derivatives of
    # order 1e-2, so the 64-cell gradient is checked to 1e-4 relative.
    rng = np.random.default_rng(7)
    cells = rng.choice(g.n, 64, replace=False)
    h = 1e-6
    fds = []
    for c in cells:
        e = np.zeros(g.n)
        e[c] = h
        fds.append((loss(A.forward(om0 + e, dts)) - loss(A.forward(om0 - e, dts))) / (2 * h))
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
        assert F.true_relres() <= 1.5e-8
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
