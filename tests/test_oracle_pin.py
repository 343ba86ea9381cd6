"""Pins the CPU oracle (oracle/ggb_oracle.c) before it is trusted as the checker.

float build ("f32"):  bit-for-bit against the golden fixtures written by the
                      reference's own compiled core, and against that core directly
                      (oracle/_ref) when it is present.
double build ("f64"): against what the mathematics fixes -- the reference's own
                      invariant tests (spherical-harmonic Laplacian, PCG accuracy,
                      curl(u) = omega, solid-body rotation and particle return), a dense
                      least-squares solve at N = 4, symmetry of K and of the V-cycle,
                      constant-field and extremum properties of the advection, a
                      closed-form single splat -- and agreement with the float build.
The planted-mistake test checks that these pins detect a wrong operator.
"""
import os

import numpy as np
import pytest

from conftest import levels_for

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- float build: bitwise

def test_f32_matches_golden_fluid(ggb, orc):
    """set_state + 5 steps of the reference (fixture) reproduced bit for bit."""
    orc.set_omp_threads(1)
    gold = np.load(os.path.join(GOLD, "fluid_N8.npz"))
    levels = levels_for(ggb, orc, 8)
    O = orc.OracleFluid("f32", levels, forcing=0.5, rotation=2.0)
    O.set_state(gold["vel0"], gold["vband"])
    assert np.array_equal(bits(O.get(0)), bits(gold["omega0"]))
    assert np.array_equal(bits(O.get(2)), bits(gold["psi0"]))
    assert np.array_equal(bits(O.get(3)), bits(gold["vel_0"]))
    assert O.cg_iters == int(gold["its0"])
    for s in range(5):
        dt = levels[0].h / max(O.max_speed(), 1e-6)
        assert dt == gold["dt"][s]
        O.step(dt)
        assert O.cg_iters == gold["its"][s]
    assert np.array_equal(bits(O.get(0)), bits(gold["omega5"]))
    assert np.array_equal(bits(O.get(2)), bits(gold["psi5"]))
    assert np.array_equal(bits(O.get(3)), bits(gold["vel5"]))


def test_f32_matches_golden_particles_render(ggb, orc):
    orc.set_omp_threads(1)
    gold = np.load(os.path.join(GOLD, "fluid_N8.npz"))
    gp = np.load(os.path.join(GOLD, "particles.npz"))
    levels = levels_for(ggb, orc, 8)
    O = orc.OracleFluid("f32", levels, forcing=0.5, rotation=2.0)
    O.set_state(gold["vel0"], gold["vband"])
    P = orc.OracleParticles("f32", 500)
    P.color_from_image(ggb.strip_to_float(ggb.banded_strip(5)), 64, 1024)
    assert np.array_equal(bits(P.col), bits(gp["col"]))
    R = orc.OracleRender("f32", 32)
    for s in range(5):
        dt = levels[0].h / max(O.max_speed(), 1e-6)
        P.advect(levels[0], O.get(3).astype(np.float32), dt)
        O.step(dt)
        R.fade(0.02)
        R.deposit(P)
        assert np.array_equal(bits(P.pos), bits(gp["pos"][s]))
    assert np.array_equal(bits(R.buf), bits(gp["buf"]))
    assert R.g == float(gp["g"])


@pytest.mark.parametrize("N,forcing,rotation,bfecc,mg", [(16, 0.5, 2.0, 1, 1), (20, 0.0, 0.0, 0, 1),
                                                         (12, 0.2, 5.0, 1, 0)])
def test_f32_matches_compiled_reference(ggb, orc, ref, N, forcing, rotation, bfecc, mg):
    levels = levels_for(ggb, orc, N)
    g = levels[0]
    rg = orc.RefGrid(N)
    vel0, vband = g.initial_velocity(seed=N, octaves=3)
    O = orc.OracleFluid("f32", levels, forcing=forcing, rotation=rotation, bfecc=bfecc, multigrid=mg,
                        cg_maxit=3000)
    RF = orc.RefFluid(rg, forcing=forcing, rotation=rotation, bfecc=bfecc, multigrid=mg, cg_maxit=3000)
    vb = vband if forcing > 0 else None
    O.set_state(vel0, vb)
    RF.set_state(vel0, vb)
    for _ in range(6):
        dt = g.h / max(RF.max_speed(), 1e-6)
        O.step(dt)
        RF.step(dt)
        assert O.cg_iters == RF.cg_iters
    assert np.array_equal(bits(O.get(0)), bits(RF.omega))
    assert np.array_equal(bits(O.get(2)), bits(RF.psi))
    assert np.array_equal(bits(O.get(3)), bits(RF.vel))
    if mg:
        for l in range(len(levels) - 1):
            pt, pw = O.transfer(l)
            rpt, rpw = RF.transfer(l)
            assert np.array_equal(pt, rpt)
            assert np.array_equal(bits(pw), bits(rpw))


def test_f32_render_matches_compiled_reference(ggb, orc, ref):
    """Render of the reference (bins, chunks, threads) == the oracle's sequential splat."""
    g = ggb.Grid(16)
    rg = orc.RefGrid(16)
    vel0, _ = g.initial_velocity(seed=2, octaves=2)
    O = orc.OracleFluid("f32", levels_for(ggb, orc, 16))
    O.set_state(vel0)
    P = orc.OracleParticles("f32", 3000)
    RP = orc.RefParticles(3000)
    img = ggb.strip_to_float(ggb.banded_strip(2))
    P.color_from_image(img, 64, 1024)
    RP.color_from_image(img, 64, 1024)
    R = orc.OracleRender("f32", 96, 1.5, 0.5)
    RR = orc.RefRender(96, 1.5, 0.5)
    for _ in range(4):
        dt = g.h / max(O.max_speed(), 1e-6)
        v = O.get(3).astype(np.float32)
        P.advect(g, v, dt)
        RP.advect(rg, v, dt)
        O.step(dt)
        R.fade(0.3)
        RR.fade(0.3)
        R.deposit(P)
        RR.deposit(RP)
    assert np.array_equal(bits(R.buf), bits(RR.buf))


# ---------------------------------------------------------------- double build: mathematics

def test_f64_laplacian_spherical_harmonics(ggb, orc):
    """reference tests/test_poisson.c:10-49: psi = xy + z/2, Lap = -6xy - z."""
    O = orc.Oracle("f64")
    prev = None
    for N in (16, 32, 64):
        g = ggb.Grid(N)
        p = g.pos.astype(np.float64)
        y = O.laplacian(g, p[:, 0] * p[:, 1] + 0.5 * p[:, 2])
        e = y - (-6 * p[:, 0] * p[:, 1] - p[:, 2])
        rms = np.sqrt(np.sum(e * e * g.area) / g.total_area)
        if N >= 64:
            assert rms < 0.05
        if prev:
            assert 1.6 < prev / rms < 2.6  # first-order convergence
        prev = rms


def test_f64_K_symmetric_and_nullspace(ggb, orc):
    g = ggb.Grid(12)
    O = orc.Oracle("f64")
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(g.n), rng.standard_normal(g.n)
    assert abs(x @ O.applyK(g, y) - y @ O.applyK(g, x)) < 1e-12 * abs(x @ O.applyK(g, x))
    assert np.max(np.abs(O.applyK(g, np.ones(g.n)))) == 0.0


def test_f64_pcg_matches_dense_solve(ggb, orc):
    """PCG (tight tolerance) vs a dense least-squares solve of K psi = -A omega at N = 4."""
    levels = levels_for(ggb, orc, 8)
    g = levels[0]
    O = orc.Oracle("f64")
    K = np.array([O.applyK(g, e) for e in np.eye(g.n)]).T
    rng = np.random.default_rng(1)
    om = rng.standard_normal(g.n)
    a = g.area.astype(np.float64)
    om -= np.sum(a * om) / np.sum(a)  # b = -A om exactly orthogonal to K's null space
    F = orc.OracleFluid("f64", levels, cg_tol=1e-10, cg_maxit=500)
    F.set(0, om)
    F.solve_psi()
    psi = F.get(2)
    b = -g.area.astype(np.float64) * om
    ref, *_ = np.linalg.lstsq(K, b, rcond=None)
    ref -= np.sum(g.area * ref) / g.total_area
    assert F.cg_relres <= 1e-10
    assert rel(psi, ref) < 1e-8


def test_f64_vcycle_symmetric(ggb, orc):
    levels = levels_for(ggb, orc, 16)
    F = orc.OracleFluid("f64", levels)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(levels[0].n), rng.standard_normal(levels[0].n)
    mx, my = F.vcycle(x), F.vcycle(y)
    assert abs(y @ mx - x @ my) < 1e-12 * abs(x @ mx)
    assert x @ mx > 0


def test_f64_poisson_and_curl(ggb, orc):
    """reference tests/test_poisson.c:51-88"""
    levels = levels_for(ggb, orc, 64)
    g = levels[0]
    O = orc.Oracle("f64")
    p = g.pos.astype(np.float64)
    F = orc.OracleFluid("f64", levels, cg_tol=1e-6, cg_maxit=5000)
    F.set(0, O.remove_mean(g, -6 * p[:, 0] * p[:, 1] - p[:, 2]))
    F.solve_psi()
    assert F.cg_relres <= 1e-6
    ex = p[:, 0] * p[:, 1] + 0.5 * p[:, 2]
    assert np.sqrt(np.sum((F.get(2) - ex) ** 2 * g.area) / np.sum(ex ** 2 * g.area)) < 0.02
    F.compute_velocity()
    om = O.curl(g, F.get(3))
    w = F.get(0)
    assert np.sqrt(np.sum((om - w) ** 2 * g.area) / np.sum(w ** 2 * g.area)) < 0.05


def test_f64_solid_body_rotation(ggb, orc):
    """reference tests/test_flow.c: u = A (y x p); particles return after one revolution."""
    levels = levels_for(ggb, orc, 64)
    g = levels[0]
    p = g.pos.astype(np.float64)
    A = 1.0
    vel0 = (A * np.cross(np.array([0.0, 1.0, 0.0]), p)).astype(np.float32)
    F = orc.OracleFluid("f64", levels, cg_tol=1e-6, cg_maxit=2000)
    F.set_state(vel0)
    v = F.get(3).reshape(-1, 3)
    vrel = np.sqrt(np.sum(np.sum((v - vel0) ** 2, 1) * g.area) / np.sum(np.sum(vel0.astype(float) ** 2, 1) * g.area))
    assert vrel < 0.01
    O = orc.Oracle("f64")
    assert np.max(np.abs(O.divergence(g, F.get(3)))) < 0.05
    ex = 2 * A * p[:, 1]
    om0 = F.get(0)
    assert np.sqrt(np.sum((om0 - ex) ** 2 * g.area) / np.sum(ex ** 2 * g.area)) < 0.01
    P = orc.OracleParticles("f64", 2000)
    start = P.pos.copy().reshape(-1, 3)
    steps = 400
    dt = 2 * np.pi / A / steps
    for _ in range(steps):
        P.advect(g, F.get(3), dt)
        F.step(dt)
    end = P.pos.reshape(-1, 3)
    ang = np.arctan2(np.linalg.norm(np.cross(end, start), axis=1), np.sum(end * start, 1))
    assert ang.max() < g.h
    om = F.get(0)
    assert np.sqrt(np.sum((om - om0) ** 2 * g.area) / np.sum(om0 ** 2 * g.area)) < 0.02


def test_f64_particle_rk2_closed_form(ggb, orc):
    """One RK2 step in exact solid-body rotation about +Y equals the rotation R_y(A dt)
    to the interpolation error of the sampled velocity (small dt)."""
    g = ggb.Grid(32)
    p = g.pos.astype(np.float64)
    A = 0.7
    vel = A * np.cross(np.array([0.0, 1.0, 0.0]), p)
    P = orc.OracleParticles("f64", 1000)
    x0 = P.pos.copy().reshape(-1, 3)
    dt = 0.01
    P.advect(g, vel.ravel(), dt)
    th = A * dt
    Ry = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    ex = x0 @ Ry.T
    assert np.max(np.linalg.norm(P.pos.reshape(-1, 3) - ex, axis=1)) < 5e-5


def test_f64_advection_constant_and_extrema(ggb, orc):
    """A constant vorticity field advects to itself; BFECC + limiter makes no new extrema."""
    levels = levels_for(ggb, orc, 16)
    g = levels[0]
    vel0, _ = g.initial_velocity(seed=7, octaves=2)
    F = orc.OracleFluid("f64", levels, cg_tol=1e-8)
    F.set_state(vel0)
    om = F.get(0)
    lo, hi = om.min(), om.max()
    for _ in range(3):
        F.step(g.h / max(F.max_speed(), 1e-6))
    om2 = F.get(0)
    # no forcing: the limited update stays inside the previous range up to the mean shift
    assert om2.min() >= lo - 1e-9 - abs(np.sum(om2 * g.area)) and om2.max() <= hi + 1e-9
    # constant field through the sampler: bilinear weights sum to one
    idx, wt = orc.Oracle("f64").taps(g, g.pos[5].astype(np.float64))
    assert abs(wt.sum() - 1.0) < 1e-15


def test_f64_single_splat_closed_form(orc):
    """One particle that does not move: one sample at its position, tent weights
    w = a c_a (1 - |x - px|/r)(1 - |y - py|/r), buffer = w c (g = 1)."""
    P = orc.OracleParticles("f64", 1)
    p = np.array([0.1, 0.05, 1.0])
    p /= np.linalg.norm(p)
    P.pos[:] = p
    P.prev[:] = p
    P.col[:] = [0.2, 0.4, 0.6, 1.0]
    R = orc.OracleRender("f64", 64, 1.0, 0.5)
    R.deposit(P)
    u, v = p[0] / p[2], p[1] / p[2]
    px = np.float32((u + 1) * 0.5 * 64 - 0.5)
    py = np.float32((1 - v) * 0.5 * 64 - 0.5)
    buf = R.buf[0]
    for y in range(64):
        for x in range(64):
            wx = max(0.0, 1 - abs(x - float(px)))
            wy = max(0.0, 1 - abs(y - float(py)))
            w = 0.5 * wx * wy
            assert abs(buf[y, x, 3] - w) < 1e-6
            assert abs(buf[y, x, 0] - 0.2 * w) < 1e-6
    assert np.count_nonzero(R.buf[1:]) == 0


def test_f64_agrees_with_f32(ggb, orc):
    levels = levels_for(ggb, orc, 16)
    g = levels[0]
    vel0, vband = g.initial_velocity(seed=11, octaves=2)
    A = orc.OracleFluid("f64", levels, forcing=0.5)
    B = orc.OracleFluid("f32", levels, forcing=0.5)
    A.set_state(vel0, vband)
    B.set_state(vel0, vband)
    dt = g.h / max(A.max_speed(), 1e-6)
    A.step(dt)
    B.step(dt)
    assert rel(A.get(0), B.get(0)) < 1e-3
    assert rel(A.get(2), B.get(2)) < 1e-3


def test_planted_mistake_is_caught(ggb, orc):
    """A Laplacian with one neighbour weight dropped fails the spherical-harmonic pin."""
    g = ggb.Grid(64)
    p = g.pos.astype(np.float64)
    x = p[:, 0] * p[:, 1] + 0.5 * p[:, 2]
    ex = -6 * p[:, 0] * p[:, 1] - p[:, 2]
    O = orc.Oracle("f64")
    good = O.laplacian(g, x)
    saved = g.w.copy()
    try:
        g.w[:, 0] = 0.0  # the mistake: drop the longest edge of every cell
        bad = O.laplacian(g, x)
    finally:
        g.w[:] = saved
    rms = lambda y: np.sqrt(np.sum((y - ex) ** 2 * g.area) / g.total_area)
    assert rms(good) < 0.05
    assert rms(bad) > 0.05


# ---------------------------------------------------------------- output path (render.c:55-66, 220-279;
# main.c:249-318) against the compiled reference

def _deposited_pair(ggb, orc, R=40, fade=0.3, steps=4, radius=1.5, alpha=0.5):
    """Oracle (f32) and reference renders after the same particle history."""
    g = ggb.Grid(16)
    rg = orc.RefGrid(16)
    vel0, _ = g.initial_velocity(seed=4, octaves=2)
    O = orc.OracleFluid("f32", levels_for(ggb, orc, 16))
    O.set_state(vel0)
    P = orc.OracleParticles("f32", 2500)
    RP = orc.RefParticles(2500)
    img = ggb.strip_to_float(ggb.banded_strip(6))
    P.color_from_image(img, 64, 1024)
    RP.color_from_image(img, 64, 1024)
    Ro = orc.OracleRender("f32", R, radius, alpha)
    Rr = orc.RefRender(R, radius, alpha)
    gs = []
    for _ in range(steps):
        dt = g.h / max(O.max_speed(), 1e-6)
        v = O.get(3).astype(np.float32)
        P.advect(g, v, dt)
        RP.advect(rg, v, dt)
        O.step(dt)
        Ro.fade(fade)
        Rr.fade(fade)
        gs.append(Rr.g)
        Ro.deposit(P)
        Rr.deposit(RP)
    return Ro, Rr, gs


def test_f32_fade_renormalisation_matches_compiled_reference(ggb, orc, ref):
    """fade 0.9 over five deposits takes g below 1e-4 (render.c:57): the buffer is scaled
    by (float) g and g reset to 1 -- the same bits as the reference, at every step."""
    Ro, Rr, gs = _deposited_pair(ggb, orc, fade=0.9, steps=5)
    assert 1.0 in gs, "the renormalisation branch did not run: %s" % gs
    assert np.array_equal(bits(Ro.buf), bits(Rr.buf))
    assert Ro.g == Rr.g
    assert np.count_nonzero(Ro.buf[..., 3]) > 100


def test_f32_output_readers_match_compiled_reference(ggb, orc, ref):
    """render_texel + render_pixel8 over every texel, with and without a background, and
    render_sample_dir at random directions, face centres, edges and corners: the oracle's
    bytes and floats equal the reference's."""
    Ro, Rr, _ = _deposited_pair(ggb, orc, fade=0.2, steps=3)
    R = Ro.R
    bg = np.array([0.25, 0.5, 0.75], np.float32)
    f8, f8bg = Ro.faces8(), Ro.faces8(bg)
    assert np.count_nonzero(f8[..., 3]) > 100
    for f in range(6):
        for y in range(R):
            for x in range(R):
                c = Rr.texel(f, x, y)
                assert np.array_equal(f8[f, y, x], Rr.pixel8(c)), (f, x, y)
                assert np.array_equal(f8bg[f, y, x], Rr.pixel8(c, bg)), (f, x, y)
    rng = np.random.default_rng(5)
    dirs = list(rng.normal(size=(3000, 3)))
    for a in (0.0, 1e-9, 0.3, 1.0 - 1e-12, 1.0):  # up to the face edges and corners
        for s in ((1, a, 0.2), (a, 1, -a), (-a, 0.4, 1), (1, 1, a), (1, -1, 1)):
            for sign in (1, -1):
                dirs.append(sign * np.array(s, float))
    for d in dirs:
        assert np.array_equal(bits(Ro.sample_dir(d)), bits(Rr.sample_dir(d))), d


def _png(path):
    from PIL import Image
    return np.asarray(Image.open(path))


def _run_reference_program(tmp, args):
    import subprocess
    exe = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ggbref_main")
    if not os.path.exists(exe):
        pytest.skip("reference program (oracle/_ref/ggbref_main) not built")
    subprocess.run([exe] + args, check=True, cwd=tmp, capture_output=True, timeout=300)


def _oracle_run_like_main(ggb, orc, img8, N, M, R, steps, seed, forcing, rotation):
    """The reference driver's set-up and step loop (main.c:383-434) on the f32 oracle."""
    levels = levels_for(ggb, orc, N)
    g = levels[0]
    vel0, vband = g.initial_velocity(seed=seed)
    O = orc.OracleFluid("f32", levels, forcing=forcing, rotation=rotation, bfecc=1, cg_tol=1e-4, cg_maxit=500)
    O.set_state(vel0, vband if forcing > 0 else None)
    P = orc.OracleParticles("f32", M)
    h, w = img8.shape[:2]
    P.color_from_image(ggb.strip_to_float(img8), w, h)
    Ro = orc.OracleRender("f32", R, 1.0, 0.35)
    for _ in range(steps):
        vmax = O.max_speed()
        dt = 1.0 * g.h / (vmax if vmax > 1e-6 else 1e-6)
        P.advect(g, O.get(3).astype(np.float32), dt)
        O.step(dt)
        Ro.fade(0.02)
        Ro.deposit(P)
    return Ro


def test_f32_outputs_match_reference_program(ggb, orc, ref, tmp_path):
    """The reference program itself (main.c + image_io.c + core, one thread) writes its
    six faces, cross net and equirectangular preview; the oracle run through the same
    set-up and steps gives the same PNG bytes, with a transparent and an opaque
    background (main.c:249-318, render.c:220-279)."""
    pytest.importorskip("PIL")
    from PIL import Image
    img8 = ggb.banded_strip(3)
    Image.fromarray(img8, "RGBA").save(str(tmp_path / "in.png"))
    N, M, R, steps, seed = 16, 3000, 48, 6, 3
    Ro = _oracle_run_like_main(ggb, orc, img8, N, M, R, steps, seed, 0.5, 2.0)
    common = ["-i", "in.png", "--grid", str(N), "--particles", str(M), "--res", str(R), "--steps", str(steps),
              "--every", str(steps), "--seed", str(seed), "--forcing", "0.5", "--rotation", "2", "--threads", "1",
              "--net", "--preview"]
    for bgarg, bg in (("transparent", None), ("4080c0", np.array([0x40, 0x80, 0xc0], np.float32) / np.float32(255))):
        out = "out_" + bgarg
        _run_reference_program(str(tmp_path), common + ["-o", out, "--bg", bgarg])
        faces = Ro.faces8(bg)
        assert np.count_nonzero(faces) > 0
        for f in range(6):
            assert np.array_equal(_png(tmp_path / out / ("gg%d.png" % f)), faces[f]), (bgarg, f)
        assert np.array_equal(_png(tmp_path / out / "gg-net.png"), Ro.net8(bg)), bgarg
        assert np.array_equal(_png(tmp_path / out / "gg-preview.png"), Ro.preview8(bg=bg)), bgarg
