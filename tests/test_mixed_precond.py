"""Pins the float64 restatement of the mixed-precision preconditioner (tests/
mixed_precond.py), the checker of the GPU's Chebyshev V-cycle:

  * with the levels' own Voronoi operators and degree 1 everywhere its Chebyshev step is
    the damped Jacobi step, and the V-cycle equals the oracle's reference V-cycle
    (fluid.c:113-146);
  * its smoother's error propagator is the Chebyshev polynomial for [1/2, 2],
    q_k(lam) = T_k((5/2 - 2 lam) / (3/2)) / T_k(5/3), evaluated through an
    eigendecomposition of D^-1 K;
  * its Galerkin coarse operators: symmetric with zero row sums, their half-precision
    weights within half a unit in the last place of the products, a 5x5 stencil on the
    face interior, the same energy as the level's own Voronoi operator on smooth fields,
    and -- what defines them -- the two-grid coarse correction I - P K_c^+ R K_f they
    give is a K_f-orthogonal projection (it is not with the Voronoi coarse operator, nor
    with a planted mistake in the product's diagonal);
  * the library's own host construction of those operators (csrc/host/galerkin.c, CPU
    code called through ctypes) equals them bit for bit, level by level;
  * the V-cycle with the default degrees is symmetric (the same polynomial before and
    after the coarse correction);
  * the exact bilinear face-interior transfer weights of the Chebyshev hierarchy are the
    reference's own taps at the exact cell centres.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import levels_for
from mixed_precond import MixedVcycle, chebyshev_hierarchy, exact_weights, galerkin, kmat


def make(ggb, orc, N, galerkin_levels=True):
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    tr = []
    for l in range(len(levels) - 1):
        pt, pw = O.transfer(l)
        tr.append((np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4)))
    return levels, O, MixedVcycle(levels, tr, galerkin_levels)


@pytest.mark.parametrize("N", [16, 32])
def test_degree_one_is_the_reference_vcycle(ggb, orc, N):
    levels, O, M = make(ggb, orc, N, galerkin_levels=False)
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


def _exact_transfer(ggb, orc, N):
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    pt, pw = O.transfer(0)
    pt, pw = exact_weights(np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4), N)
    nf, nc = levels[0].n, levels[1].n
    P = sp.csr_matrix((pw.ravel(), (np.repeat(np.arange(nf), 4), pt.ravel())), shape=(nf, nc))
    return levels, P


@pytest.mark.parametrize("N", [32, 64])
def test_galerkin_weights(ggb, orc, N):
    levels, P = _exact_transfer(ggb, orc, N)
    K = kmat(levels[0])
    G = galerkin(P.T.tocsr(), K, P)
    Gx = galerkin(P.T.tocsr(), K, P, half=False)
    # symmetric with zero row sums (some couplings are positive: see the spectrum test)
    assert abs(G - G.T).max() == 0.0
    assert np.abs(np.asarray(G.sum(axis=1))).max() <= 1e-12 * G.diagonal().max()
    # half-precision storage: each weight within half an ulp of the product's
    W, Wx = (sp.diags(G.diagonal()) - G).tocoo(), (sp.diags(Gx.diagonal()) - Gx).tocsr()
    ex = np.asarray(Wx[W.row, W.col]).ravel()
    assert np.all(np.abs(W.data - ex) <= 2.0 ** -11 * np.abs(ex) + 2.0 ** -25)
    # a 5x5 stencil on the face interior
    Nc = N // 2
    i = np.arange(G.shape[0])
    f, y, x = i // (Nc * Nc), (i // Nc) % Nc, i % Nc
    Gc = G.tocoo()
    inner = (x[Gc.row] >= 2) & (x[Gc.row] <= Nc - 3) & (y[Gc.row] >= 2) & (y[Gc.row] <= Nc - 3)
    d = Gc.col[inner] - Gc.row[inner]
    assert np.all(f[Gc.col[inner]] == f[Gc.row[inner]])
    assert np.all(np.abs(np.round(d / Nc)) <= 2) and np.all(np.abs(d - np.round(d / Nc) * Nc) <= 2)
    # on smooth fields its energy is the coarse grid's own Voronoi operator's
    Kc = kmat(levels[1])
    pos = np.asarray(levels[1].pos, np.float64).reshape(-1, 3)
    for v in (pos[:, 0], pos[:, 1] * pos[:, 2], pos[:, 2] ** 2):
        v = v - v.mean()
        assert abs((v @ (G @ v)) / (v @ (Kc @ v)) - 1.0) < 0.01


@pytest.mark.parametrize("N", [32, 64])
def test_galerkin_levels_spectrum_inside_smoothing_interval(ggb, orc, N):
    """The Chebyshev smoother damps [1/2, 2] of the spectrum of D^-1 K.  On level 0 the
    weights are non-negative and each row of D^-1 W sums to one, so Gershgorin places the
    spectrum in [0, 2].  The Galerkin levels have negative weights (positive couplings,
    up to about a quarter of the diagonal summed over a row), where Gershgorin allows up
    to ~2.5: the bound is checked on their spectra instead (largest eigenvalue 1.55-1.72
    on hierarchies from N = 64 and 256)."""
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    tr = []
    for l in range(len(levels) - 1):
        pt, pw = O.transfer(l)
        pt, pw = np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4)
        tr.append(exact_weights(pt, pw, levels[l].N) if chebyshev_hierarchy(N) else (pt, pw))
    M = MixedVcycle(levels, tr, galerkin_levels=True)
    W0 = (sp.diags(M.K[0].diagonal()) - M.K[0]).tocoo()
    assert W0.data.min() >= 0.0
    negative = 0
    for l in range(1, len(levels)):
        K = M.K[l]
        d = K.diagonal()
        W = (sp.diags(d) - K).tocoo()
        negative += int((W.data < 0).sum())
        s = 1.0 / np.sqrt(d)
        S = (sp.diags(s) @ K @ sp.diags(s)).toarray()
        lam = np.linalg.eigvalsh(S)
        print("level %d (%d cells): lambda(D^-1 K) in [%.2e, %.4f]" % (l, K.shape[0], lam[0], lam[-1]))
        assert lam[0] > -1e-10 and lam[-1] < 2.0
    assert negative > 0


def test_galerkin_coarse_correction_is_a_projection(ggb, orc):
    levels, P = _exact_transfer(ggb, orc, 16)
    K = kmat(levels[0]).toarray()
    Pd = P.toarray()
    R = Pd.T
    rng = np.random.default_rng(4)
    X = rng.standard_normal((K.shape[0], 3))

    def defects(Kc):
        E = np.eye(K.shape[0]) - Pd @ np.linalg.pinv(Kc) @ R @ K
        proj = np.linalg.norm(E @ (E @ X) - E @ X) / np.linalg.norm(E @ X)
        sym = np.linalg.norm(K @ E - (K @ E).T) / np.linalg.norm(K @ E)
        return proj, sym

    proj, sym = defects(galerkin(sp.csr_matrix(R), sp.csr_matrix(K), P, half=False).toarray())
    assert proj < 1e-7 and sym < 1e-7
    # the rediscretised coarse operator, and a product with a planted mistake, are not
    # (K E stays symmetric for any symmetric coarse operator: only the projection tells)
    proj, sym = defects(kmat(levels[1]).toarray())
    assert proj > 1e-3
    Kbad = R @ K @ Pd
    Kbad = Kbad + 0.05 * np.diag(np.diag(Kbad))  # a planted mistake in the diagonal
    proj, sym = defects(Kbad)
    assert proj > 1e-3


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


class _GalOp(C.Structure):
    # the library's host Galerkin operator (csrc/host/ggb_host.h: GalOp)
    _fields_ = [("n", C.c_int), ("N", C.c_int), ("nbr", C.POINTER(C.c_int)), ("w", C.POINTER(C.c_double)),
                ("wh", C.POINTER(C.c_ushort)), ("nn", C.POINTER(C.c_ubyte)), ("reg", C.POINTER(C.c_ubyte)),
                ("dinv", C.POINTER(C.c_float))]


@pytest.mark.parametrize("N", [32, 64])
def test_library_galerkin_operators_match_restatement(ggb, orc, N):
    """The library's host construction of the mixed hierarchy's coarse operators
    (csrc/host/galerkin.c, CPU code) gives, level by level, exactly the restatement's
    half-precision Galerkin products, with a regular cell's slots at the fixed 5x5 offsets
    and dinv = 1 / sum of its weights."""
    L = ggb.lib()
    L.ggb_galerkin.restype = C.c_int
    L.ggb_galop_free.restype = None
    levels = levels_for(ggb, orc, N)
    O = orc.OracleFluid("f64", levels)
    K = kmat(levels[0])
    keep, prev = [], None
    for l in range(len(levels) - 1):
        pt, pw = O.transfer(l)
        pt, pw = exact_weights(np.asarray(pt).reshape(-1, 4), np.asarray(pw).reshape(-1, 4), levels[l].N)
        pt = np.ascontiguousarray(pt, np.int32)
        pw = np.ascontiguousarray(pw, np.float32)
        nf, nc = levels[l].n, levels[l + 1].n
        fine = np.repeat(np.arange(nf), 4)
        P = sp.csr_matrix((pw.ravel().astype(np.float64), (fine, pt.ravel())), shape=(nf, nc))
        # R = P^T as rows with fine indices ascending
        order = np.lexsort((fine, pt.ravel()))
        ridx = np.ascontiguousarray(fine[order], np.int32)
        rw = np.ascontiguousarray(pw.ravel()[order], np.float32)
        rstart = np.zeros(nc + 1, np.int32)
        np.add.at(rstart, pt.ravel()[order] + 1, 1)
        rstart = np.cumsum(rstart).astype(np.int32)
        out = _GalOp()
        ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
        rc = L.ggb_galerkin(C.byref(levels[0].c), C.byref(prev) if prev is not None else None,
                            pt.ctypes.data_as(ip), pw.ctypes.data_as(fp), rstart.ctypes.data_as(ip),
                            ridx.ctypes.data_as(ip), rw.ctypes.data_as(fp), levels[l + 1].N, C.byref(out))
        assert rc == 0
        keep.append(out)
        n = out.n
        nbr = np.ctypeslib.as_array(out.nbr, (n * 24,)).reshape(n, 24)
        w = np.ctypeslib.as_array(out.w, (n * 24,)).reshape(n, 24)
        reg = np.ctypeslib.as_array(out.reg, (n,)).astype(bool)
        dinv = np.ctypeslib.as_array(out.dinv, (n,))
        W = sp.csr_matrix((w.ravel(), (np.repeat(np.arange(n), 24), nbr.ravel())), shape=(n, n))
        W.setdiag(0)
        W.eliminate_zeros()
        Klib = (sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W).tocsr()
        Kref = galerkin(P.T.tocsr(), K, P)
        assert abs(Klib - Kref).max() == 0.0
        assert np.array_equal(dinv, np.float32(1.0 / Kref.diagonal()))
        Nc = levels[l + 1].N
        i = np.nonzero(reg)[0]
        for k in range(24):
            q = k if k < 12 else k + 1
            assert np.all(nbr[i, k] == i + (q // 5 - 2) * Nc + (q % 5 - 2))
        K, prev = Kref, out
    for g in keep:
        L.ggb_galop_free(C.byref(g))
