"""The mixed-precision preconditioner on the GPU (csrc/kern/mgt_k.hip: Chebyshev
smoothing of degree 4 on level 0 and 2 on the Galerkin coarse levels below, fused tile
passes, float storage) against its float64 restatement (tests/mixed_precond.py, pinned by
test_mixed_precond.py), and its effect on the PCG:

  * the V-cycle applied through fluid_vcycle_f64 matches the restatement (with the exact
    bilinear face-interior transfer weights of the Chebyshev hierarchy) to float rounding
    (2e-6 relative), at sizes with interior tiles and at sizes made only of seam tiles;
    grids whose levels are not multiples of the 16-cell tile keep the reference V(1,1)
    cycle in float and its transfer (the restatement with degree 1);
  * every tile through the graph tables gives the same bits as the interior tiles'
    index arithmetic;
  * the V-cycle is symmetric and positive to float rounding;
  * on a developed flow at tol 1e-8 the mixed solve needs fewer iterations than the FP64
    one and still reaches a true FP64 relative residual of 1e-8.
"""
import os

import numpy as np
import pytest

from conftest import levels_for
from mixed_precond import MixedVcycle, chebyshev_hierarchy, exact_weights

pytestmark = pytest.mark.gpu


def restatement(ggb, orc, N):
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    tr = []
    for l in range(len(levels) - 1):
        pt, pw = O.transfer(l)
        pt, pw = np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4)
        if chebyshev_hierarchy(N):
            pt, pw = exact_weights(pt, pw, levels[l].N)
        tr.append((pt, pw))
    return levels, MixedVcycle(levels, tr, galerkin_levels=chebyshev_hierarchy(N))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_mixed_vcycle_matches_restatement(ggb, orc, gpu, N):
    levels, M = restatement(ggb, orc, N)
    F = ggb.Fluid(levels[0], precision=ggb.PREC_MIXED)
    rng = np.random.default_rng(N)
    for _ in range(2):
        b = rng.standard_normal(levels[0].n)
        e = rel(F.vcycle(b), M(b))
        print("N=%d mixed V-cycle vs restatement: %.2e" % (N, e))
        assert e < 2e-6


def test_mixed_vcycle_untiled_grid_keeps_reference_cycle(ggb, orc, gpu):
    levels, M = restatement(ggb, orc, 24)
    F = ggb.Fluid(levels[0], precision=ggb.PREC_MIXED)
    b = np.random.default_rng(5).standard_normal(levels[0].n)
    assert rel(F.vcycle(b), M(b, degrees=lambda l: 1)) < 2e-6


def test_graph_tables_match_interior_tiles(ggb, orc, gpu):
    levels = levels_for(ggb, orc, 128)
    b = np.random.default_rng(9).standard_normal(levels[0].n)
    F1 = ggb.Fluid(levels[0], precision=ggb.PREC_MIXED)
    z1 = F1.vcycle(b)
    os.environ["GGB_TILES_GRAPH"] = "1"
    try:
        F2 = ggb.Fluid(levels[0], precision=ggb.PREC_MIXED)
    finally:
        del os.environ["GGB_TILES_GRAPH"]
    z2 = F2.vcycle(b)
    assert np.array_equal(z1.view(np.uint64), z2.view(np.uint64))


@pytest.mark.parametrize("N", [32, 64])
def test_mixed_vcycle_symmetric(ggb, orc, gpu, N):
    levels = levels_for(ggb, orc, N)
    F = ggb.Fluid(levels[0], precision=ggb.PREC_MIXED)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(levels[0].n), rng.standard_normal(levels[0].n)
    mx, my = F.vcycle(x), F.vcycle(y)
    assert abs(mx @ y - x @ my) <= 1e-6 * np.linalg.norm(mx) * np.linalg.norm(y)
    assert mx @ x > 0


def test_mixed_iterations_and_residual(ggb, orc, gpu):
    """N=128 with forcing 0.5 and rotation 10 (the benchmarked flow): 20 steps in each
    precision from the same state with the same dt; the mixed solve takes fewer CG
    iterations per step and every solve reaches the true residual."""
    g = ggb.Grid(128)
    vel0, vband = g.initial_velocity(seed=1, octaves=4)
    its = {}
    for prec in (ggb.PREC_FP64, ggb.PREC_MIXED):
        F = ggb.Fluid(g, forcing=0.5, rotation=10.0, cg_tol=1e-8, precision=prec)
        F.set_state(vel0, vband)
        its[prec] = []
        for s in range(20):
            F.step(0.5 * g.h)
            its[prec].append(F.cg_iters)
            assert F.true_relres() <= 1e-8
    print("CG iterations per step: FP64 %s, mixed %s" % (its[ggb.PREC_FP64], its[ggb.PREC_MIXED]))
    assert np.mean(its[ggb.PREC_MIXED]) <= 0.75 * np.mean(its[ggb.PREC_FP64])


def test_post_solve_passes_follow_a_short_batch(ggb, gpu):
    """fluid_step enqueues the passes after its solve (psi's mean, the velocity, max |u|^2)
    before it reads the loop flag; when the batch, predicted from the previous solve, was
    too short they run again after the next one.  After very short steps (a nearly exact
    warm start, few iterations) psi is set to zero: the next solve starts cold and needs
    more iterations than predicted; the velocity and the max speed afterwards are those of
    the final psi
    (u = p x sum_k gc_k (psi_j - psi_i), restated on the host in float64) -- a pass left on
    an unconverged iterate would differ by at least a CG iteration's change."""
    g = ggb.Grid(64)
    vel0, vband = g.initial_velocity(seed=3, octaves=4)
    F = ggb.Fluid(g, forcing=0.5, rotation=10.0, cg_tol=1e-8, precision=ggb.PREC_MIXED)
    F.set_state(vel0, vband)
    F.step(0.5 * g.h)
    for _ in range(4):
        F.step(1e-4 * g.h)
    its0 = F.cg_iters
    F.set(ggb.FIELD_PSI, np.zeros(g.n))
    F.step(0.5 * g.h)
    print("CG iterations: %d, then %d" % (its0, F.cg_iters))
    assert F.cg_iters > its0
    psi = F.get(ggb.FIELD_PSI)
    vel = F.get(ggb.FIELD_VEL).reshape(-1, 3)
    p = g.pos.astype(np.float64)
    nnb = g.nnb.astype(int)
    d = np.zeros((g.n, 3))
    for k in range(8):
        m = k < nnb
        d[m] += g.gc[m, k].astype(np.float64) * (psi[g.nbr[m, k]] - psi[m])[:, None]
    u = np.cross(p, d)
    assert np.linalg.norm(vel - u) <= 1e-10 * np.linalg.norm(u)
    vmax = np.sqrt(np.max(np.sum(vel ** 2, axis=1)))
    assert abs(F.max_speed() - vmax) <= 1e-12 * vmax


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_zero_vorticity_gives_zero_streamfunction(ggb, gpu, prec):
    """A zero right-hand side (b = -A omega = 0: omega, its band, rotation and forcing all
    zero) stops the solve before its first iteration and sets psi to zero, as the
    reference's PCG does for bb == 0 (fluid.c:306-311); the zeroing runs in the pass
    after the solve (with the psi mean), and the extrapolated warm start of the mixed
    path, formed from a nonzero psi, must not survive it."""
    g = ggb.Grid(64)
    P = ggb.PREC_FP64 if prec == "fp64" else ggb.PREC_MIXED
    F = ggb.Fluid(g, forcing=0.0, rotation=0.0, cg_tol=1e-8, precision=P)
    z = np.zeros(g.n)
    F.set(ggb.FIELD_OMEGA, z)
    F.set(ggb.FIELD_OMEGA_BAND, z)
    F.set(ggb.FIELD_VEL, np.zeros(3 * g.n))
    F.set(ggb.FIELD_PSI, np.random.default_rng(4).standard_normal(g.n))
    for _ in range(3):
        F.step(0.5 * g.h)
        assert F.cg_iters == 0
        assert np.all(F.get(ggb.FIELD_PSI) == 0.0)
        assert np.all(F.get(ggb.FIELD_VEL) == 0.0)
        assert F.max_speed() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_converged_start_takes_no_iteration(ggb, gpu, prec):
    """A step so short that the warm start already meets the tolerance: the solve stops
    before its first iteration (the reference tests the residual first, fluid.c:311-321),
    psi stays the warm start, the true relative residual is within the tolerance, and the
    velocity and max speed are those of that psi."""
    g = ggb.Grid(64)
    vel0, vband = g.initial_velocity(seed=5, octaves=4)
    P = ggb.PREC_FP64 if prec == "fp64" else ggb.PREC_MIXED
    F = ggb.Fluid(g, forcing=0.5, rotation=10.0, cg_tol=1e-8, precision=P)
    F.set_state(vel0, vband)
    for _ in range(3):
        F.step(0.5 * g.h)
    F.step(1e-9 * g.h)
    assert F.cg_iters == 0
    assert F.true_relres() <= 1e-8
    psi = F.get(ggb.FIELD_PSI)
    vel = F.get(ggb.FIELD_VEL).reshape(-1, 3)
    p = g.pos.astype(np.float64)
    nnb = g.nnb.astype(int)
    d = np.zeros((g.n, 3))
    for k in range(8):
        m = k < nnb
        d[m] += g.gc[m, k].astype(np.float64) * (psi[g.nbr[m, k]] - psi[m])[:, None]
    u = np.cross(p, d)
    assert np.linalg.norm(vel - u) <= 1e-10 * np.linalg.norm(u)
    vmax = np.sqrt(np.max(np.sum(vel ** 2, axis=1)))
    assert abs(F.max_speed() - vmax) <= 1e-12 * vmax
