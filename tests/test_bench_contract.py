"""bench.py's command-line contract and workload definition, without a GPU: one GPU
only (a larger --gpus exits with status 2 and says so), the default timed window is the
N=512 workload's whole run (2000 steps after 5 warm-up steps), the roofline is the draw's
by default, the workload is the one the metric is quoted on, the draw's algorithmic bytes
follow the render contract (prev, pos and colour per particle in float, texels read and
written), the fill's the per-pass model, and the fluid step's aggregate is the sum of its
passes (every V-cycle's and CG iteration's passes, plus the fixed ones)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_default_window_is_the_whole_run():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'add_argument("--steps", type=int, default=2000)' in src
    assert 'add_argument("--warmup", type=int, default=5)' in src


def test_roofline_kernel_defaults_to_the_draw(bench):
    # the draw has the most device time over the whole run; the warm-up's early steps
    # do not decide it unless asked ("auto")
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'add_argument("--roofline-kernel", default="k_render_draw"' in src
    assert "k_render_draw" in bench.PROBES


def test_workload_is_the_quoted_configuration(bench):
    c2 = bench.CONFIGS["c2"]
    assert (c2["N"], c2["P"], c2["R"]) == (512, 4_000_000, 2048)
    assert (c2["forcing"], c2["rotation"], c2["cg_tol"]) == (0.5, 10.0, 1e-8)
    assert c2["precision"] == bench.ggb.PREC_MIXED
    c4 = bench.CONFIGS["c4"]
    assert (c4["N"], c4["P"], c4["R"], c4["octaves"]) == (1024, 16_000_000, 4096, 6)


def test_byte_model(bench):
    n, P, E, T = 1_572_864, 4_000_000, 5_000_000, 19_000_000
    # draw: 40 B per particle (prev, pos, colour in float) + 16 B read and 16 B written per
    # touched texel
    assert bench.kernel_bytes("k_render_draw", n, P, E, T, 8, 4, 4) == 40 * P + 32 * T
    # fill: endpoints and sample count always; colour, index and the copy only when reordering
    plain = bench.kernel_bytes("k_render_fill", n, P, E, T, 8, 4, 4, 0.0)
    reord = bench.kernel_bytes("k_render_fill", n, P, E, T, 8, 4, 4, 1.0)
    assert plain == P * 50 + E * 4
    assert reord - plain == P * 88
    with pytest.raises(KeyError):
        bench.kernel_bytes("no_such_kernel", n, P, E, T, 8, 4, 4)


def test_fluid_step_byte_model(bench):
    ggb = bench.ggb
    b6, parts6 = bench.fluid_step_bytes(512, 6, 7, ggb.PREC_MIXED, True)
    b7, parts7 = bench.fluid_step_bytes(512, 7, 8, ggb.PREC_MIXED, True)
    n = 6 * 512 * 512
    # one more iteration of the mixed PCG adds one fused update + down pass, one level-0 up
    # pass (which also forms A z: no separate Ap pass) and one pass over the Galerkin coarse
    # levels (126 B per cell, 48 of them the half-precision weights) and the dense bottom --
    # about 180 B per fine cell
    extra = (b7 - b6) / n
    assert 150 < extra < 210, extra
    assert "Ap" not in parts6
    coarse = n * (1 / 4 + 1 / 16 + 1 / 64 + 1 / 256)
    assert extra * n == pytest.approx(bench.kernel_bytes("k_cg_update_down", n, 0, 0, 0, 8, 4, 4, cheb=True)
                                      + bench.kernel_bytes("mg_up(level 0)", n, 0, 0, 0, 8, 4, 4, cheb=True)
                                      + 126 * coarse + 4 * 1536 ** 2 + 8 * 1536)
    # the rediscretised coarse levels (no weights per cell)
    b6v, _ = bench.fluid_step_bytes(512, 6, 7, ggb.PREC_MIXED, True, galerkin=False)
    b7v, _ = bench.fluid_step_bytes(512, 7, 8, ggb.PREC_MIXED, True, galerkin=False)
    assert (b7v - b6v) - (b7 - b6) == pytest.approx((46 - 126) * coarse)
    assert sum(parts6.values()) == int(b6) or abs(sum(parts6.values()) - b6) < len(parts6)
    # FP64 storage moves more per V-cycle, and its PCG has the Ap pass
    b6d, parts6d = bench.fluid_step_bytes(512, 6, 7, ggb.PREC_FP64, False)
    assert b6d > b6
    assert parts6d["Ap"] == 6 * bench.kernel_bytes("k_cg_Ap", n, 0, 0, 0, 8, 8, 8)

