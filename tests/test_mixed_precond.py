"""Pins the float64 restatement of the mixed-precision preconditioner (tests/
mixed_precond.py), the checker of the GPU's Chebyshev V-cycle:

  * with degree 1 everywhere its Chebyshev step is the damped Jacobi step, and the
    V-cycle equals the oracle's reference V-cycle (fluid.c:113-146);
  * its smoother's error propagator is the Chebyshev polynomial for [1/2, 2],
    q_k(lam) = T_k((5/2 - 2 lam) / (3/2)) / T_k(5/3), evaluated through an
    eigendecomposition of D^-1 K;
  * the V-cycle with the default degrees is symmetric (the same polynomial before and
    after the coarse correction).
"""
import numpy as np
import pytest

from conftest import levels_for
from mixed_precond import MixedVcycle, kmat


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
