"""The oracle's tape adjoint of the fluid step, pinned by central finite differences.

The adjoint (oracle/ggb_oracle.c, orc_adj_*) is the reverse-mode derivative of the
restated fluid_step (reference src/fluid.c:397-478) with a prescribed dt schedule.  It is
checked against central differences of the same forward map along random directions
of the initial vorticity and in the forcing strength.  The map is piecewise smooth (the
limiter's active set and the bilinear tap cells switch), so the step is small enough not
to cross a switch; the Poisson solves run at tol 1e-10, where the warm-started PCG still
converges in FP64.
"""
import numpy as np
import pytest

from conftest import levels_for


def setup_case(ggb, orc, N, bfecc, forcing=0.5, rotation=2.0, seed=3):
    levels = levels_for(ggb, orc, N)
    g = levels[0]
    O = orc.Oracle("f64")
    vel0, vband = g.initial_velocity(seed=seed, octaves=3)
    om0 = O.curl(g, vel0.astype(np.float64))
    omb = O.curl(g, vband.astype(np.float64))
    omb -= np.sum(g.area * omb) / g.total_area
    F = orc.OracleFluid("f64", levels, forcing=forcing, rotation=rotation, bfecc=bfecc, cg_tol=1e-10,
                        cg_maxit=2000)
    F.set(1, omb)
    rng = np.random.default_rng(seed)
    target = 0.5 * om0[rng.permutation(g.n)]
    return g, F, om0, target, rng


def loss_and_grad(F, g, om0, target, dts):
    A = g.area.astype(np.float64)
    wT = F.adj_forward(om0, dts)
    L = 0.5 * np.sum(A * (wT - target) ** 2)
    g0, gk = F.adj_backward(A * (wT - target))
    return L, g0, gk


def fd(F, g, om0, target, dts, d, eps):
    A = g.area.astype(np.float64)
    Lp = 0.5 * np.sum(A * (F.adj_forward(om0 + eps * d, dts) - target) ** 2)
    Lm = 0.5 * np.sum(A * (F.adj_forward(om0 - eps * d, dts) - target) ** 2)
    return (Lp - Lm) / (2 * eps)


@pytest.mark.parametrize("N,nsteps,bfecc", [(8, 4, 1), (16, 4, 0)])
def test_oracle_adjoint_matches_fd(ggb, orc, N, nsteps, bfecc):
    g, F, om0, target, rng = setup_case(ggb, orc, N, bfecc)
    dts = np.full(nsteps, 0.6 * g.h)
    L, g0, gk = loss_and_grad(F, g, om0, target, dts)
    assert np.all(np.isfinite(g0)) and np.isfinite(gk)
    for _ in range(2):
        d = rng.standard_normal(g.n)
        ad = float(g0 @ d)
        err = min(abs(fd(F, g, om0, target, dts, d, eps) - ad) for eps in (1e-5, 1e-6))
        assert err <= 1e-6 * abs(ad), (err, ad)
    # forcing strength
    k0 = 0.5
    errs = []
    for eps in (1e-5, 1e-6):
        F.set_forcing(k0 + eps)
        Lp = 0.5 * np.sum(g.area * (F.adj_forward(om0, dts) - target) ** 2)
        F.set_forcing(k0 - eps)
        Lm = 0.5 * np.sum(g.area * (F.adj_forward(om0, dts) - target) ** 2)
        errs.append(abs((Lp - Lm) / (2 * eps) - gk))
    F.set_forcing(k0)
    assert min(errs) <= 1e-6 * abs(gk), (errs, gk)
