"""Writes the golden fixtures in tests/golden/ from the reference's own numerical core
(oracle/_ref/libggbref.so, compiled from the reference sources by oracle/Makefile),
run on one OpenMP thread.  Inputs are seeded and built by this repository's library.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  geometry_N8.npz  -- cs_init(8): pos, area, nnb, nbr, w, elen, gc (first nnb slots)
  fluid_N8.npz     -- initial velocity (seed 5), then omega/psi/vel after set_state and
                      after 5 steps (forcing 0.5, rotation 2, BFECC, tol 1e-4), dt per
                      step, CG iterations per step
  particles.npz    -- 500 Fibonacci particles coloured from the seeded strip and advected
                      5 steps in the fluid above; a 32^2 render buffer of those steps
"""
import json
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gaseous-gusty-ball_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ggb  # noqa: E402
import ggb_oracle as O  # noqa: E402


def main():
    O.set_omp_threads(1)
    N = 8
    r = O.RefGrid(N)
    k = np.arange(8)[None, :] < r.nnb[:, None]
    np.savez(os.path.join(HERE, "geometry_N8.npz"), pos=r.pos, area=r.area, nnb=r.nnb,
             nbr=np.where(k, r.nbr, -1), w=np.where(k, r.w, 0), elen=np.where(k, r.elen, 0),
             gc=np.where(k[..., None], r.gc, 0), total_area=r.total_area)

    g = ggb.Grid(N)
    vel0, vband = g.initial_velocity(seed=5, octaves=2)
    F = O.RefFluid(r, forcing=0.5, rotation=2.0)
    F.set_state(vel0, vband)
    out = dict(vel0=vel0, vband=vband, omega0=F.omega.copy(), psi0=F.psi.copy(), vel_0=F.vel.copy(),
               its0=F.cg_iters)
    P = O.RefParticles(500)
    img = ggb.strip_to_float(ggb.banded_strip(5))
    P.color_from_image(img, 64, 1024)
    col = P.col.copy()
    R = O.RefRender(32)
    dts, its, pos = [], [], []
    for _ in range(5):
        dt = r.h / max(F.max_speed(), 1e-6)
        P.advect(r, F.vel, dt)
        F.step(dt)
        R.fade(0.02)
        R.deposit(P)
        dts.append(dt)
        its.append(F.cg_iters)
        pos.append(P.pos.copy())
    out.update(omega5=F.omega.copy(), psi5=F.psi.copy(), vel5=F.vel.copy(), dt=np.array(dts),
               its=np.array(its))
    np.savez(os.path.join(HERE, "fluid_N8.npz"), **out)
    np.savez(os.path.join(HERE, "particles.npz"), col=col, pos=np.array(pos), buf=R.buf.copy(), g=R.g)
    sums = {}
    for f in ("geometry_N8.npz", "fluid_N8.npz", "particles.npz"):
        sums[f] = hashlib.sha256(open(os.path.join(HERE, f), "rb").read()).hexdigest()
    json.dump(sums, open(os.path.join(HERE, "SHA256SUMS.json"), "w"), indent=1)
    print("wrote", ", ".join(sums))


if __name__ == "__main__":
    main()
