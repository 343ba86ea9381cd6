"""Host geometry of the library (cs_init, sampling, seam ghost table) against the
reference's compiled cs_init (bitwise), the golden fixture, the reference's own
invariants (tests/test_cubesphere.c), and an independent textbook spherical Voronoi
diagram (scipy) at small N.  CPU only."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def stencil_view(g):
    k = np.arange(8)[None, :] < g.nnb[:, None]
    return k


def test_geometry_matches_golden(ggb):
    gold = np.load(os.path.join(HERE, "golden", "geometry_N8.npz"))
    g = ggb.Grid(8)
    k = stencil_view(g)
    assert np.array_equal(g.pos.view(np.uint32), gold["pos"].view(np.uint32))
    assert np.array_equal(g.area.view(np.uint32), gold["area"].view(np.uint32))
    assert np.array_equal(g.nnb, gold["nnb"])
    assert np.array_equal(g.nbr[k], gold["nbr"][k])
    assert np.array_equal(g.w[k].view(np.uint32), gold["w"][k].view(np.uint32))
    assert np.array_equal(g.elen[k].view(np.uint32), gold["elen"][k].view(np.uint32))
    assert np.array_equal(g.gc[k].view(np.uint32), gold["gc"][k].view(np.uint32))
    assert g.total_area == float(gold["total_area"])


@pytest.mark.parametrize("N", [4, 6, 8, 12, 20, 32, 64, 127])
def test_geometry_matches_compiled_reference(ggb, orc, ref, N):
    g = ggb.Grid(N)
    r = orc.RefGrid(N)
    k = stencil_view(g)
    assert np.array_equal(g.nnb, r.nnb)
    for name in ("pos", "area"):
        assert np.array_equal(getattr(g, name).view(np.uint32), getattr(r, name).view(np.uint32)), name
    assert np.array_equal(g.nbr[k], r.nbr[k])
    for name in ("w", "elen"):
        assert np.array_equal(getattr(g, name)[k].view(np.uint32), getattr(r, name)[k].view(np.uint32)), name
    assert np.array_equal(g.gc[k].view(np.uint32), r.gc[k].view(np.uint32))
    assert g.total_area == r.total_area
    assert (g.c.asym_count, g.c.open_cells) == (r.asym_count, r.open_cells)


def test_reference_invariants(ggb):
    """tests/test_cubesphere.c:30-75 at N = 64"""
    g = ggb.Grid(64)
    assert g.c.asym_count == 0 and g.c.open_cells == 0
    k = stencil_view(g)
    nb = np.where(k, g.nbr, -1)
    assert not np.any(nb == np.arange(g.n)[:, None])  # no self links
    assert g.nnb.min() >= 3
    assert abs(g.total_area - 4 * np.pi) / (4 * np.pi) < 1e-4
    # reciprocity: j in nbr(i) <=> i in nbr(j), with equal weights
    rows = np.repeat(np.arange(g.n), g.nnb)
    cols = g.nbr[k]
    w = g.w[k]
    fwd = dict(zip(zip(rows, cols), w))
    assert all(fwd[(j, i)] == wij for (i, j), wij in fwd.items())
    # centres map back to their cell
    assert all(g.cell_of(g.pos[c]) == c for c in range(0, g.n, 97))


def test_locate_round_trip(ggb, orc):
    """cs_locate -> face angles -> direction round trip (test_cubesphere.c:30-44)"""
    g = ggb.Grid(64)
    O = orc.Oracle("f64")
    rng = np.random.default_rng(12345)
    d = rng.standard_normal((2000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    for p in d[:200]:
        idx, wt = O.taps(g, p)
        assert abs(wt.sum() - 1) < 1e-15 and np.all(wt >= -1e-15)
        assert np.all((idx >= 0) & (idx < g.n))
        # bilinear reconstruction of the position field is close to p
        q = (wt[:, None] * g.pos[idx]).sum(0)
        assert np.linalg.norm(q - p) < 2 * g.h


def test_ghost_table_matches_cell_of(ggb, orc):
    """every ghost entry is the real cell holding the extended-face centre"""
    N = 16
    g = ggb.Grid(N)
    O = orc.Oracle("f64")
    fc = np.array([[0, 0, 1], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], float)
    fr = np.array([[1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [1, 0, 0], [1, 0, 0]], float)
    fu = np.array([[0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], float)
    for f in range(6):
        for j in range(-1, N + 1):
            for i in (-1, 0, N - 1, N):
                a = -np.pi / 4 + (i + 0.5) * g.h
                b = -np.pi / 4 + (j + 0.5) * g.h
                d = fc[f] + fr[f] * np.tan(a) + fu[f] * np.tan(b)
                d /= np.linalg.norm(d)
                want = (f * N + j) * N + i if 0 <= i < N and 0 <= j < N else O.cell_of(g, d)
                assert g.ghost[f, j + 1, i + 1] == want


@pytest.mark.parametrize("N", [4, 8])
def test_against_textbook_spherical_voronoi(ggb, N):
    """scipy's SphericalVoronoi of the (float) centres: same neighbour sets, edge lengths
    and areas to float precision."""
    sv_mod = pytest.importorskip("scipy.spatial")
    g = ggb.Grid(N)
    pts = g.pos.astype(np.float64)
    pts /= np.linalg.norm(pts, axis=1)[:, None]
    sv = sv_mod.SphericalVoronoi(pts, radius=1.0, center=np.zeros(3))
    sv.sort_vertices_of_regions()
    areas = sv.calculate_areas()
    assert np.max(np.abs(areas - g.area) / areas) < 1e-5
    for c in range(g.n):
        reg = sv.regions[c]
        verts = sv.vertices[reg]
        edges = {}
        for a in range(len(reg)):
            v0, v1 = verts[a], verts[(a + 1) % len(reg)]
            ln = np.arctan2(np.linalg.norm(np.cross(v0, v1)), v0 @ v1)
            # the neighbour across this edge: the other region sharing both vertices
            shared = [r for r in range(g.n) if r != c and reg[a] in sv.regions[r]
                      and reg[(a + 1) % len(reg)] in sv.regions[r]]
            if len(shared) == 1 and ln >= 1e-4 * g.h:
                edges[shared[0]] = ln
        mine = {int(g.nbr[c, k]): float(g.elen[c, k]) for k in range(g.nnb[c])}
        assert set(mine) == set(edges), c
        for j, ln in edges.items():
            assert abs(mine[j] - ln) < 1e-5 * max(ln, g.h)


def decode_stencil(N, n, rec, cw, xn):
    """Neighbour and weight tables [n, 8] from the compressed stencil."""
    mags = np.array([(c >> 2) * N + (c & 3) - N - 1 for c in range(16)], np.int64)
    idx = np.arange(n, dtype=np.int64)
    cls = (rec[:, 0].astype(np.int64) & 0x7FFFFFFF)
    irr = rec[:, 0] < 0
    code = np.where(irr, 0, rec[:, 1].astype(np.int64) & 0xFFFFFFFF)
    nib = (code[:, None] >> (4 * np.arange(8))) & 15
    nbr = idx[:, None] + mags[nib]
    nbr[irr] = xn[rec[irr, 1]]
    return nbr, cw[cls]


@pytest.mark.parametrize("N", [8, 16, 33, 64, 128])
def test_stencil_compression_round_trip(ggb, N):
    """The compressed stencil (weight classes + offset codes) decodes to the ELL table
    exactly, slot for slot, with (self, 0) padding; the classes really are shared."""
    g = ggb.Grid(N)
    rec, cw, xn = g.stencil_compressed()
    nbr, w = decode_stencil(N, g.n, rec, cw, xn)
    nnb = g.nnb.astype(int)
    slot = np.arange(8)[None, :]
    real = slot < nnb[:, None]
    assert np.array_equal(nbr[real], g.nbr[real])
    assert np.array_equal(w[real].view(np.uint32), g.w[real].view(np.uint32))
    assert np.all(nbr[~real] == np.broadcast_to(np.arange(g.n)[:, None], nbr.shape)[~real])
    assert np.all(w[~real] == 0)
    # explicit rows only along the face seams (about 4 per face row), and every class
    # shared by many cells
    assert len(xn) <= 5 * g.n / N
    assert len(cw) <= g.n / 10 if N >= 16 else len(cw) < g.n
