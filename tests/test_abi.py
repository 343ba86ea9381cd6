"""The drop-in boundary: libggb.so loads and exports every function the public headers
(include/*.h) declare; the headers compile as C; a reference-style C driver compiles
and links against the library.  No device calls (CPU only)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "gaseous-gusty-ball_amd", "lib")

DECL = re.compile(r"^(?!static|typedef|#|\s|\*|/|extern \"C\"|\})[A-Za-z_][\w \*]*?\b(\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INC, h)).read()
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libggb.so")],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_every_declared_function_is_exported(ggb):
    decl = declared_functions()
    assert len(decl) > 40
    missing = sorted(decl - exported_symbols())
    assert not missing, missing
    for name in decl:
        getattr(ggb.lib(), name)  # resolvable through the loader as well


def test_face_basis_tables_exported():
    syms = exported_symbols()
    for name in ("cs_face_center", "cs_face_right", "cs_face_up"):
        assert name in syms


def test_binding_lists_every_public_symbol(ggb):
    assert declared_functions() <= set(ggb.PUBLIC_SYMBOLS) | {"cs_face_center", "cs_face_right", "cs_face_up"}


DRIVER = r"""
#include <stdio.h>
#include "cubesphere.h"
#include "fluid.h"
#include "particles.h"
#include "render.h"
#include "noise.h"
#include "ggb_ext.h"
#include "ggb_inputs.h"
int main(int argc, char **argv)
{
	(void) argv;
	CubeSphere cs; Fluid F; Particles P; Render R;
	if (argc < 2) return 0; /* link check only: never touches the device */
	cs_init(&cs, 16);
	FluidParams fp = { 0.5, 2.0, 1, 1e-4, 500, 1 };
	fluid_init(&F, &cs, &fp);
	particles_init(&P, 100);
	render_init(&R, 64, 1.0f, 0.35f);
	double dt = cs.h / fluid_max_speed(&F);
	particles_advect(&P, &cs, F.vel, dt);
	fluid_step(&F, dt);
	render_fade(&R, 0.02);
	render_deposit(&R, &P);
	float c[4]; unsigned char px[4];
	render_texel(&R, 0, 0, 0, c);
	render_pixel8(c, NULL, px);
	printf("%d %g\n", F.cg_iters, F.cg_relres);
	render_free(&R); particles_free(&P); fluid_free(&F); cs_free(&cs);
	return 0;
}
"""


def test_reference_style_driver_compiles_and_links(tmp_path):
    src = tmp_path / "drv.c"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    # the reference builds with -Wall -Wextra (warnings allowed); FluidParams gained a
    # trailing field, which a six-value initialiser leaves at its FP64 default
    subprocess.run(["gcc", "-std=gnu11", "-Wall", "-Wextra", "-Werror", "-Wno-missing-field-initializers", "-I" + INC, str(src), "-o", str(exe),
                    "-L" + LIBDIR, "-lggb", "-Wl,-rpath," + LIBDIR, "-lm"], check=True)
    # runs without touching the device (no argument)
    subprocess.run([str(exe)], check=True)
