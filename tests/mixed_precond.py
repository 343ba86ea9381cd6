"""Restatement of the mixed-precision preconditioner in float64 numpy -- test
infrastructure: the checker of the GPU's Chebyshev V-cycle (csrc/kern/mgt_k.hip).

The V-cycle follows the reference's (fluid.c:113-146: zero start, pre-smooth, residual,
restriction R = P^T, coarse correction, prolongation, post-smooth, 60 damped Jacobi
sweeps on the coarsest level) with the changes the mixed mode makes:
  * the smoother is the Chebyshev iteration on D^-1 K for the interval [1/2, 2], of
    degree 4 on level 0 and 2 below;
  * the coarse levels carry Galerkin operators K_(l+1) = R_l K_l P_l instead of their own
    Voronoi Laplacians, in the form (K x)_i = sum_j w_ij (x_i - x_j) with the symmetrised
    off-diagonal weights w_ij = -(K_ij + K_ji) / 2 rounded to half precision (the
    library's storage), each level formed from the rounded level above.
The library tabulates the cycle below its first level of at most 1536 cells as a dense
matrix; that changes only the rounding.  With the rediscretised levels and degree 1
everywhere the Chebyshev step is the damped Jacobi step, so the restatement is then the
reference V-cycle, which the CPU tests check against the oracle's.
"""
import numpy as np
import scipy.sparse as sp


def chebyshev_hierarchy(N):
    """Whether the library builds the Chebyshev V-cycle for a grid of face size N: a first
    coarse level of at most 1536 cells (the dense bottom), every level above it a multiple
    of the 16-cell tile (csrc/host/fluid.c, cheb_eligible)."""
    nlev = 1
    while (N >> nlev) >= 4 and (N >> (nlev - 1)) % 2 == 0:
        nlev += 1
    L = next((l for l in range(1, nlev) if 6 * (N >> l) ** 2 <= 1536), -1)
    return L >= 1 and all((N >> l) % 16 == 0 for l in range(L))


def exact_weights(taps, weights, Nf):
    """The Chebyshev hierarchy's transfer: fine cells (f, x, y) with 1 <= x, y <= Nf - 2 whose
    taps follow the bilinear pattern (coarse ((x - 1) >> 1, (y - 1) >> 1) and the cells right,
    up and up-right of it) take the exact bilinear weights in equiangular coordinates -- the
    centre sits a quarter cell from its nearer tap column, so 3/4 there and 1/4 on the other
    (odd x: the first column), the same for rows -- instead of those evaluated at the
    float-rounded centres; other cells keep theirs."""
    Nc = Nf // 2
    taps = np.asarray(taps).reshape(-1, 4)
    w = np.array(weights, np.float64).reshape(-1, 4)
    i = np.arange(taps.shape[0])
    f, y, x = i // (Nf * Nf), (i // Nf) % Nf, i % Nf
    inner = (x >= 1) & (x <= Nf - 2) & (y >= 1) & (y <= Nf - 2)
    b = (f * Nc + ((y - 1) >> 1)) * Nc + ((x - 1) >> 1)
    pat = inner & (taps[:, 0] == b) & (taps[:, 1] == b + 1) & (taps[:, 2] == b + Nc) & (taps[:, 3] == b + Nc + 1)
    ax = np.where(x & 1, 0.75, 0.25)
    ay = np.where(y & 1, 0.75, 0.25)
    ex = np.stack([ax * ay, (1 - ax) * ay, ax * (1 - ay), (1 - ax) * (1 - ay)], 1)
    w[pat] = ex[pat]
    return taps, w


def kmat(g):
    """K x = sum_k w_k (x_i - x_nb(k)) as a sparse matrix (float weights, float64 sums)."""
    n = g.n
    nbr = g.nbr.reshape(n, 8)
    w = g.w.reshape(n, 8).astype(np.float64)
    rows, cols, vals = [], [], []
    for k in range(8):
        i = np.nonzero(k < g.nnb)[0]
        rows += [i, i]
        cols += [i, nbr[i, k]]
        vals += [w[i, k], -w[i, k]]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def cheb_coefficients(deg):
    """(cd_k, cr_k): d_k = cd_k d_{k-1} + cr_k D^-1 r_k for [1/2, 2] (sigma 5/3, delta 3/4)."""
    sigma, delta = 5.0 / 3.0, 0.75
    rho = 1.0 / sigma
    out = [(0.0, 1.0 / 1.25)]
    for _ in range(1, deg):
        rn = 1.0 / (2 * sigma - rho)
        out.append((rn * rho, 2 * rn / delta))
        rho = rn
    return out


def galerkin(R, K, P, half=True):
    """R K P as (K x)_i = sum_j w_ij (x_i - x_j): symmetrised off-diagonal weights
    -(G_ij + G_ji) / 2, rounded to half precision when half (the library's storage)."""
    G = (R @ K @ P).tocoo()
    off = G.row != G.col
    W = sp.csr_matrix((-G.data[off], (G.row[off], G.col[off])), shape=G.shape)
    W = ((W + W.T) * 0.5).tocsr()
    if half:
        W.data = W.data.astype(np.float16).astype(np.float64)
    W.eliminate_zeros()
    return (sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W).tocsr()


class MixedVcycle:
    def __init__(self, levels, transfers, galerkin_levels=True, deg0=4):
        """levels: ggb.Grid per level; transfers[l] = (taps [nf, 4], weights [nf, 4]).
        galerkin_levels: coarse operators R K P (else each level's own Voronoi operator,
        degree 3 below level 0 and deg0 = 2: the hierarchy without them)."""
        self.levels = levels
        self.galerkin = galerkin_levels
        self.deg0 = deg0 if galerkin_levels else 2
        self.P = []
        for l, (pt, pw) in enumerate(transfers):
            nf, nc = levels[l].n, levels[l + 1].n
            self.P.append(sp.csr_matrix((np.asarray(pw, np.float64).reshape(nf, 4).ravel(),
                                         (np.repeat(np.arange(nf), 4), np.asarray(pt).reshape(nf, 4).ravel())),
                                        shape=(nf, nc)))
        self.R = [P.T.tocsr() for P in self.P]
        self.K = [kmat(levels[0])]
        for l in range(1, len(levels)):
            self.K.append(galerkin(self.R[l - 1], self.K[l - 1], self.P[l - 1]) if galerkin_levels
                          else kmat(levels[l]))
        self.dinv = [np.float32(1.0 / K.diagonal()).astype(np.float64) for K in self.K]

    def smooth(self, l, b, x, deg):
        """deg Chebyshev steps from x on K x = b."""
        d = None
        for cd, cr in cheb_coefficients(deg):
            r = self.dinv[l] * (b - self.K[l] @ x)
            d = cr * r if d is None else cd * d + cr * r
            x = x + d
        return x

    def jacobi(self, l, b, x, sweeps):
        for _ in range(sweeps):
            x = x + 0.8 * self.dinv[l] * (b - self.K[l] @ x)
        return x

    def __call__(self, b, l=0, degrees=None):
        """z = M b; degrees(l) gives the smoothing degree (default deg0 on level 0, below it
        2 with the Galerkin levels, 3 without)."""
        if degrees is None:
            deg = self.deg0 if l == 0 else (2 if self.galerkin else 3)
        else:
            deg = degrees(l)
        n = self.levels[l].n
        if l == len(self.levels) - 1:
            return self.jacobi(l, b, np.zeros(n), 60)
        x = self.smooth(l, b, np.zeros(n), deg)
        x = x + self.P[l] @ self(self.R[l] @ (b - self.K[l] @ x), l + 1, degrees)
        return self.smooth(l, b, x, deg)
