"""Pins the float64 restatement of the mixed-precision preconditioner (tests/
mixed_precond.py), the checker of the GPU's Chebyshev V-cycle:

  * with degree 1 everywhere its Chebyshev step is the damped Jacobi step, and the
    V-cycle equals the oracle's reference V-cycle (fluid.c:113-146);
  * its smoother's error propagator is the Chebyshev polynomial for [1/2, 2],
    q_k(lam) = T_k((5/2 - 2 lam) / (3/2)) / T_k(5/3), evaluated through an
    eigendecomposition of D^-1 K;
  * the V-cycle with the default degrees is symmetric (the same polynomial before and
    after the coarse correction);
  * the exact bilinear face-interior transfer weights of the Chebyshev hierarchy are the
    reference's own taps at the exact cell centres.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import levels_for
from mixed_precond import MixedVcycle, chebyshev_hierarchy, exact_weights, kmat


def make(ggb, orc, N):
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    tr = []
    for l in range(len(levels) - 1):
        pt, pw = O.transfer(l)
        tr.append((np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4)))
    return levels, O, MixedVcycle(levels, tr)


@pytest.mark.parametrize("N", [16, 32])
def test_degree_one_is_the_reference_vcycle(ggb, orc, N):
    levels, O, M = make(ggb, orc, N)
    b = np.random.default_rng(N).standard_normal(levels[0].n)
    a, ref = M(b, degrees=lambda l: 1), O.vcycle(b)
    assert np.linalg.norm(a - ref) <= 1e-12 * np.linalg.norm(ref)
    # the default degrees change it
    assert np.linalg.norm(M(b) - ref) > 1e-3 * np.linalg.norm(ref)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_smoother_is_the_chebyshev_polynomial(ggb, orc, deg):
    levels, O, M = make(ggb, orc, 8)
    K = kmat(levels[0]).toarray()
    dinv = M.dinv[0]
    s = np.sqrt(dinv)
    lam, V = np.linalg.eigh(s[:, None] * K * s[None, :])  # D^-1/2 K D^-1/2
    assert lam.min() > -1e-12 and lam.max() < 2.0
    def t(k, x):  # Chebyshev polynomial T_k by its recurrence
        a, b = np.ones_like(x), x
        for _ in range(k - 1):
            a, b = b, 2 * x * b - a
        return b if k >= 1 else a

    q = t(deg, (2.5 - 2 * lam) / 1.5) / t(deg, np.array(5.0 / 3.0))
    e = np.random.default_rng(deg).standard_normal(levels[0].n)
    err = e - M.smooth(0, K @ e, np.zeros_like(e), deg)
    ex = (V @ (q * (V.T @ (e / s)))) * s  # q(D^-1 K) e
    assert np.linalg.norm(err - ex) <= 1e-10 * np.linalg.norm(ex)
    # on [1/2, 2] the polynomial stays within 1/T_k(5/3)
    hi = lam >= 0.5
    assert np.max(np.abs(q[hi])) <= 1.0 / t(deg, np.array(5.0 / 3.0)) + 1e-12


@pytest.mark.parametrize("N", [16, 32])
def test_vcycle_symmetric(ggb, orc, N):
    levels, O, M = make(ggb, orc, N)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(levels[0].n), rng.standard_normal(levels[0].n)
    a, b = M(x) @ y, x @ M(y)
    assert abs(a - b) <= 1e-12 * np.linalg.norm(M(x)) * np.linalg.norm(y)
    assert M(x) @ x > 0


def _ref_centre(orc):
    f = orc.rlib("ref").cs_face_ab_to_dir
    f.restype = orc.RVec3
    f.argtypes = [C.c_int, C.c_double, C.c_double]
    return f


def test_exact_weights_are_the_reference_taps_at_exact_centres(ggb, orc, ref):
    """The Chebyshev hierarchy's face-interior transfer weights (exact_weights, the same
    formula as csrc/host/fluid.c and mgt_k.hip) equal the reference's own cs_taps
    (cubesphere.c) evaluated at the exact double-precision cell centres (its
    cs_face_ab_to_dir at the equiangular cell midpoints, before the rounding to float
    positions) -- so tap order and the 3/4 - 1/4 assignment are the reference's -- and
    differ from the taps at the float positions by that rounding only."""
    Nf, Nc = 64, 32
    centre = _ref_centre(orc)
    rc = orc.RefGrid(Nc)
    g = ggb.Grid(Nf)
    h = (np.pi / 2) / Nf
    rng = np.random.default_rng(7)
    cells = rng.choice(6 * Nf * Nf, 400, replace=False)
    taps = np.zeros((6 * Nf * Nf, 4), np.int64)
    wts = np.zeros((6 * Nf * Nf, 4))
    exact_at = {}
    for i in cells:
        f, y, x = i // (Nf * Nf), (i // Nf) % Nf, i % Nf
        d = centre(int(f), -np.pi / 4 + (x + 0.5) * h, -np.pi / 4 + (y + 0.5) * h)
        idx, wt = rc.taps((d.x, d.y, d.z))
        exact_at[i] = (idx, wt)
        taps[i], wts[i] = rc.taps(g.pos[i].astype(np.float64))
    _, snapped = exact_weights(taps, wts, Nf)
    checked = 0
    for i, (idx, wt) in exact_at.items():
        x, y = i % Nf, (i // Nf) % Nf
        if not (1 <= x <= Nf - 2 and 1 <= y <= Nf - 2):
            assert np.array_equal(snapped[i], wts[i])  # seam cells keep their weights
            continue
        assert np.array_equal(idx, taps[i])
        assert np.abs(snapped[i] - wt).max() < 1e-12
        assert 0 < np.abs(wts[i] - wt).max() < 5e-5
        checked += 1
    assert checked > 300


def test_chebyshev_hierarchy_sizes():
    """Grids with the tiled Chebyshev V-cycle: every level above the dense one (the first
    of at most 1536 cells) a multiple of 16 wide."""
    assert [N for N in (16, 24, 32, 48, 64, 96, 128, 256, 512, 1024) if chebyshev_hierarchy(N)] == \
        [16, 32, 64, 128, 256, 512, 1024]
