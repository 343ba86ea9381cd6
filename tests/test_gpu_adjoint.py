"""GPU discrete adjoint (fluid_adjoint_* ABI, adj_k.hip) against the oracle's FP64 tape
adjoint, against central finite differences of the GPU forward, across checkpoint
intervals, and through the torch.autograd harness (adjoint.py)."""
import numpy as np
import pytest

from conftest import levels_for

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def case(ggb, orc, N, seed=3):
    levels = levels_for(ggb, orc, N)
    g = levels[0]
    O = orc.Oracle("f64")
    vel0, vband = g.initial_velocity(seed=seed, octaves=3)
    om0 = O.curl(g, vel0.astype(np.float64))
    omb = O.curl(g, vband.astype(np.float64))
    omb -= np.sum(g.area * omb) / g.total_area
    rng = np.random.default_rng(seed)
    target = 0.5 * om0[rng.permutation(g.n)]
    return levels, g, om0, omb, target, rng


def gpu_fluid(ggb, g, omb, bfecc=1, precision=None, forcing=0.5):
    F = ggb.Fluid(g, forcing=forcing, rotation=2.0, bfecc=bfecc, cg_tol=1e-10, cg_maxit=2000,
                  precision=ggb.PREC_FP64 if precision is None else precision)
    F.set(ggb.FIELD_OMEGA_BAND, omb)
    return F


def loss(g, w, target):
    return 0.5 * float(np.sum(g.area * (w - target) ** 2))


@pytest.mark.parametrize("bfecc", [1, 0])
def test_gpu_adjoint_matches_oracle_tape(ggb, orc, gpu, bfecc):
    N, nsteps = 16, 6
    levels, g, om0, omb, target, rng = case(ggb, orc, N)
    dts = np.full(nsteps, 0.6 * g.h)
    F = gpu_fluid(ggb, g, omb, bfecc)
    A = ggb.FluidAdjoint(F, checkpoint_every=4)   # two segments: 4 + 2 steps
    wT = A.forward(om0, dts)
    g0, gk = A.backward(g.area * (wT - target))
    O = orc.OracleFluid("f64", levels, forcing=0.5, rotation=2.0, bfecc=bfecc, cg_tol=1e-10, cg_maxit=2000)
    O.set(1, omb)
    wTo = O.adj_forward(om0, dts)
    g0o, gko = O.adj_backward(g.area * (wTo - target))
    print("omega_T rel %.2e  grad rel %.2e  dL/dk %.12e vs %.12e" % (rel(wT, wTo), rel(g0, g0o), gk, gko))
    assert rel(wT, wTo) < 1e-9
    assert rel(g0, g0o) < 1e-7
    assert abs(gk - gko) <= 1e-7 * abs(gko)


def test_gpu_adjoint_matches_fd(ggb, orc, gpu):
    N, nsteps = 16, 6
    levels, g, om0, omb, target, rng = case(ggb, orc, N, seed=5)
    dts = np.full(nsteps, 0.6 * g.h)
    F = gpu_fluid(ggb, g, omb)
    A = ggb.FluidAdjoint(F, checkpoint_every=3)
    wT = A.forward(om0, dts)
    g0, gk = A.backward(g.area * (wT - target))
    for _ in range(2):
        d = rng.standard_normal(g.n)
        ad = float(g0 @ d)
        errs = []
        for eps in (1e-5, 1e-6):
            fd = (loss(g, A.forward(om0 + eps * d, dts), target) -
                  loss(g, A.forward(om0 - eps * d, dts), target)) / (2 * eps)
            errs.append(abs(fd - ad))
        print("direction: adjoint %.12e  rel err %.2e" % (ad, min(errs) / abs(ad)))
        assert min(errs) <= 1e-6 * abs(ad)
    errs = []
    for eps in (1e-5, 1e-6):
        F.set_forcing(0.5 + eps)
        lp = loss(g, A.forward(om0, dts), target)
        F.set_forcing(0.5 - eps)
        lm = loss(g, A.forward(om0, dts), target)
        errs.append(abs((lp - lm) / (2 * eps) - gk))
    F.set_forcing(0.5)
    print("forcing: adjoint %.12e  rel err %.2e" % (gk, min(errs) / abs(gk)))
    assert min(errs) <= 1e-6 * abs(gk)


def test_gpu_adjoint_checkpoint_interval(ggb, orc, gpu):
    """Recompute-from-checkpoint reproduces the stored forward: the gradient does not
    depend on the interval, bit for bit (the transposes sum in a fixed order)."""
    levels, g, om0, omb, target, rng = case(ggb, orc, 16, seed=7)
    dts = np.full(7, 0.6 * g.h)
    out = []
    for K in (1, 3, 10):
        F = gpu_fluid(ggb, g, omb)
        A = ggb.FluidAdjoint(F, checkpoint_every=K)
        wT = A.forward(om0, dts)
        out.append(A.backward(g.area * (wT - target)))
        A.free()
        F.free()
    for g0, gk in out[1:]:
        assert np.array_equal(g0.view(np.uint64), out[0][0].view(np.uint64))
        assert gk == out[0][1]


def test_gpu_adjoint_mixed_precision(ggb, orc, gpu):
    """Mixed mode (FP32 V-cycle inside FP64 CG) solves to the same tolerance, so its
    gradient matches the FP64 one to about the solver tolerance."""
    levels, g, om0, omb, target, rng = case(ggb, orc, 16, seed=9)
    dts = np.full(5, 0.6 * g.h)
    res = []
    for prec in (ggb.PREC_FP64, ggb.PREC_MIXED):
        F = gpu_fluid(ggb, g, omb, precision=prec)
        A = ggb.FluidAdjoint(F, checkpoint_every=2)
        wT = A.forward(om0, dts)
        res.append(A.backward(g.area * (wT - target)))
    print("mixed vs fp64 gradient rel %.2e" % rel(res[1][0], res[0][0]))
    assert rel(res[1][0], res[0][0]) < 1e-6
    assert abs(res[1][1] - res[0][1]) <= 1e-6 * abs(res[0][1])


def test_torch_harness_gradient_and_descent(ggb, orc, gpu):
    torch = pytest.importorskip("torch")
    import adjoint
    levels, g, om0, omb, target, rng = case(ggb, orc, 16, seed=11)
    dts = np.full(5, 0.6 * g.h)
    F = gpu_fluid(ggb, g, omb)
    rc = adjoint.RolloutContext(F, dts, checkpoint_every=2)
    area = torch.from_numpy(g.area.astype(np.float64))
    tgt = torch.from_numpy(target)
    w0 = torch.from_numpy(om0.copy()).requires_grad_(True)
    k = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    L = adjoint.area_loss(adjoint.FluidRollout.apply(w0, k, rc), tgt, area)
    L.backward()
    # the same gradient through the plain ctypes calls
    A = ggb.FluidAdjoint(F, checkpoint_every=5)
    F.set_forcing(0.5)
    wT = A.forward(om0, dts)
    g0, gk = A.backward(g.area * (wT - target))
    assert rel(w0.grad.numpy(), g0) < 1e-12
    assert abs(k.grad.item() - gk) <= 1e-12 * abs(gk)
    # gradient descent with backtracking on omega0 lowers the loss (dt stays frozen, so
    # an over-long step can leave the stable range; it is then halved)
    def evaluate(w):
        wt = torch.from_numpy(w).requires_grad_(True)
        Li = adjoint.area_loss(adjoint.FluidRollout.apply(wt, k.detach(), rc), tgt, area)
        Li.backward()
        return Li.item(), wt.grad.numpy().copy()

    losses = [L.item()]
    w, grad = om0.copy(), g0
    lr = 0.1 * losses[0] / float(grad @ grad)
    for _ in range(3):
        for _ in range(8):
            Ln, gn = evaluate(w - lr * grad)
            if np.isfinite(Ln) and Ln < losses[-1]:
                break
            lr *= 0.5
        assert Ln < losses[-1]
        w, grad = w - lr * grad, gn
        losses.append(Ln)
    print("losses", losses)
    assert losses[-1] < 0.9 * losses[0]


def test_torch_harness_two_rollouts_one_context(ggb, orc, gpu):
    """Two rollouts on one context before their backward passes: each backward returns
    its own rollout's gradient (the library keeps the last forward's checkpoints only),
    and after a backward the fluid holds the rollout's initial state with its
    velocity."""
    torch = pytest.importorskip("torch")
    import adjoint
    levels, g, om0, omb, target, rng = case(ggb, orc, 16, seed=12)
    dts = np.full(4, 0.6 * g.h)
    F = gpu_fluid(ggb, g, omb)
    rc = adjoint.RolloutContext(F, dts, checkpoint_every=2)
    area = torch.from_numpy(g.area.astype(np.float64))
    tgt = torch.from_numpy(target)
    k = torch.tensor(0.5, dtype=torch.float64)
    wa = torch.from_numpy(om0.copy()).requires_grad_(True)
    wb = torch.from_numpy(0.7 * om0[::-1].copy()).requires_grad_(True)
    La = adjoint.area_loss(adjoint.FluidRollout.apply(wa, k, rc), tgt, area)
    Lb = adjoint.area_loss(adjoint.FluidRollout.apply(wb, k, rc), tgt, area)
    (La + Lb).backward()
    for w0, grad in ((wa, wa.grad), (wb, wb.grad)):
        A = ggb.FluidAdjoint(F, checkpoint_every=4)
        F.set_forcing(0.5)
        w = w0.detach().numpy().copy()
        wT = A.forward(w, dts)
        g0, _ = A.backward(g.area * (wT - target))
        assert rel(grad.numpy(), g0) < 1e-12
        # post-backward state: omega0 after the mean removal, and its velocity
        om = F.get(ggb.FIELD_OMEGA)
        ex = w - np.sum(g.area * w) / g.total_area
        assert rel(om, ex) < 1e-12
        vel = F.get(ggb.FIELD_VEL).reshape(-1, 3)
        assert abs(F.max_speed() - np.max(np.linalg.norm(vel, axis=1))) <= 1e-12 * F.max_speed()
        A.free()


def test_adjoint_iteration_schedule(ggb, orc, gpu):
    """fluid_adjoint_iterations reports each solve's CG count; prescribing the same counts
    reproduces the rollout bit for bit, and the backward pass replays them (the same
    gradient, bit for bit)."""
    levels, g, om0, omb, target, rng = case(ggb, orc, 16, seed=13)
    dts = np.full(6, 0.6 * g.h)
    F = gpu_fluid(ggb, g, omb)
    A = ggb.FluidAdjoint(F, checkpoint_every=3)
    w1 = A.forward(om0, dts)
    its = A.iterations()
    assert len(its) == 7 and np.all(its > 0)
    g1, k1 = A.backward(g.area * (w1 - target))
    A.fix_iterations(its)
    w2 = A.forward(om0, dts)
    assert np.array_equal(A.iterations(), its)
    assert np.array_equal(w1.view(np.uint64), w2.view(np.uint64))
    g2, k2 = A.backward(g.area * (w2 - target))
    # the transposes add each target's entries in source order: the gradient repeats
    assert np.array_equal(g2.view(np.uint64), g1.view(np.uint64)) and k2 == k1
    # two more iterations everywhere: a slightly different (better converged) rollout
    A.fix_iterations(its + 2)
    w3 = A.forward(om0, dts)
    assert np.array_equal(A.iterations(), its + 2)
    assert 0 < rel(w3, w1) < 1e-8
    with pytest.raises(RuntimeError):
        A.fix_iterations(its[:3])
        A.forward(om0, dts)
    A.fix_iterations(None)
    A.free()
