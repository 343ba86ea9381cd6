"""The reference's own callers, unmodified, over libggb -- the drop-in boundary.

oracle/Makefile compiles the reference's four test programs (tests/test_*.c) and its
driver (src/main.c + src/image_io.c) where they lie, against include/, and links them
with libggb.so (oracle/_ref/dropin_*):

  * CPU: test_noise and test_cubesphere (host-only entry points) pass;
  * GPU: test_poisson and test_flow pass.  test_flow reads F.omega and P.pos after its
    step loop, so it runs with the host mirrors refreshed by every step
    (GGB_SYNC_MIRRORS=1); without that its two post-loop checks would read the
    starting values and pass vacuously;
  * GPU: the driver's PNG outputs (six faces, cross net, equirectangular preview) from
    the same command line as the reference program (oracle/_ref/ggbref_main, its own
    float core) agree within 1 LSB on >= 99.9 % of their bytes -- the drop-in computes
    in FP64, the reference in float.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def _exe(name):
    path = os.path.join(REFBIN, name)
    if not os.path.exists(path):
        pytest.skip("%s not built (oracle/Makefile needs the reference tree)" % name)
    return path


def _run(name, env=None, cwd=None, args=(), timeout=600):
    e = dict(os.environ)
    e.update(env or {})
    print("running %s %s" % (name, " ".join(args)), flush=True)
    p = subprocess.run([_exe(name)] + list(args), env=e, cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    return p


@pytest.mark.parametrize("prog", ["dropin_test_noise", "dropin_test_cubesphere"])
def test_reference_host_tests_through_libggb(prog):
    p = _run(prog)
    assert p.returncode == 0
    assert p.stdout.rstrip().endswith("OK")


@pytest.mark.gpu
def test_reference_poisson_through_libggb(gpu):
    p = _run("dropin_test_poisson")
    assert p.returncode == 0
    assert p.stdout.rstrip().endswith("test_poisson: OK")


@pytest.mark.gpu
def test_reference_flow_through_libggb(gpu):
    p = _run("dropin_test_flow", env={"GGB_SYNC_MIRRORS": "1"})
    assert p.returncode == 0
    out = p.stdout
    assert out.rstrip().endswith("test_flow: OK")
    # the post-loop checks saw the evolved state: the particles moved and came back
    # (a stale mirror would report exactly zero error)
    line = [ln for ln in out.splitlines() if ln.startswith("particles after one revolution")][0]
    err = float(line.split("error")[1].split()[0])
    assert 0 < err < 0.0245
    line = [ln for ln in out.splitlines() if ln.startswith("omega after one revolution")][0]
    assert float(line.split("change")[1].split()[0]) > 0


@pytest.mark.gpu
def test_reference_driver_through_libggb(gpu, tmp_path):
    pytest.importorskip("PIL")
    from PIL import Image

    import ggb
    img8 = ggb.banded_strip(2)
    Image.fromarray(img8, "RGBA").save(str(tmp_path / "in.png"))
    args = ["-i", "in.png", "--grid", "32", "--particles", "20000", "--res", "96", "--steps", "20",
            "--every", "20", "--seed", "2", "--forcing", "0.5", "--rotation", "3", "--net", "--preview",
            "--bg", "203040"]
    for name, out in (("ggbref_main", "ref"), ("dropin_main", "gpu")):
        p = _run(name, cwd=str(tmp_path), args=args + ["-o", out])
        assert p.returncode == 0, name
    files = ["gg%d.png" % f for f in range(6)] + ["gg-net.png", "gg-preview.png"]
    for fn in files:
        a = np.asarray(Image.open(tmp_path / "gpu" / fn)).astype(int)
        b = np.asarray(Image.open(tmp_path / "ref" / fn)).astype(int)
        assert a.shape == b.shape, fn
        d = np.abs(a - b)
        print("%s: %d of %d bytes differ, max %d" % (fn, np.count_nonzero(d), d.size, d.max()))
        assert np.count_nonzero(d > 1) <= 0.001 * d.size, fn
        assert np.ptp(b) > 0, fn  # trails drawn over the background
