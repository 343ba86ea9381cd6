"""ctypes binding of libggb.so (the drop-in C ABI) for tests, benchmarks and the
verification harness.

Mirrors the C structs of include/cubesphere.h, fluid.h, particles.h and render.h and
exposes thin Python wrappers whose methods are the C entry points.  The library is the
product; this module only calls it.  Loading fails loudly when the built library is
missing (build with `make -C gaseous-gusty-ball_amd lib` or __graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GGB_LIB selects another build of the library (for comparing build variants)
LIB_PATH = os.environ.get("GGB_LIB") or os.path.join(HERE, "lib", "libggb.so")

PREC_FP64, PREC_MIXED, PREC_REF32 = 0, 1, 2
FIELD_OMEGA, FIELD_OMEGA_BAND, FIELD_PSI, FIELD_VEL, FIELD_DEP_B, FIELD_DEP_F = range(6)
CS_MAXNB = 8


class CCubeSphere(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ncells", C.c_int), ("h", C.c_double),
        ("pos", C.POINTER(C.c_float)), ("area", C.POINTER(C.c_float)),
        ("nnb", C.POINTER(C.c_ubyte)), ("nbr", C.POINTER(C.c_int)),
        ("w", C.POINTER(C.c_float)), ("elen", C.POINTER(C.c_float)),
        ("gc", C.POINTER(C.c_float)), ("total_area", C.c_double),
        ("asym_count", C.c_int), ("open_cells", C.c_int),
        ("ghost", C.POINTER(C.c_int)), ("dev", C.c_void_p),
    ]


class CVec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class CFluidParams(C.Structure):
    _fields_ = [
        ("forcing", C.c_double), ("rotation", C.c_double), ("bfecc", C.c_int),
        ("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("multigrid", C.c_int),
        ("precision", C.c_int),
    ]


class CFluid(C.Structure):
    _fields_ = [
        ("cs", C.POINTER(CCubeSphere)), ("prm", CFluidParams), ("nlev", C.c_int),
        ("omega", C.POINTER(C.c_float)), ("omega_band", C.POINTER(C.c_float)),
        ("psi", C.POINTER(C.c_float)), ("vel", C.POINTER(C.c_float)),
        ("cg_iters", C.c_int), ("cg_relres", C.c_double), ("dev", C.c_void_p),
    ]


class CParticles(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("pos", C.POINTER(C.c_float)), ("prev", C.POINTER(C.c_float)),
        ("col", C.POINTER(C.c_float)), ("dev", C.c_void_p),
    ]


class CRender(C.Structure):
    _fields_ = [
        ("R", C.c_int), ("buf", C.POINTER(C.c_float)), ("g", C.c_double),
        ("radius", C.c_float), ("alpha", C.c_float),
        ("strip_h", C.c_int), ("nstrips", C.c_int), ("nbins", C.c_int),
        ("dev", C.c_void_p),
    ]


# Every function the public headers declare (include/*.h), for the export check.
PUBLIC_SYMBOLS = [
    # cubesphere.h
    "cs_init", "cs_free", "cs_face_of", "cs_face_ab_to_dir", "cs_locate", "cs_cell_of",
    "cs_taps", "cs_sample_scalar", "cs_sample_vec", "cs_face_center", "cs_face_right",
    "cs_face_up",
    # fluid.h
    "fluid_init", "fluid_free", "fluid_curl", "fluid_divergence", "fluid_laplacian",
    "fluid_remove_mean", "fluid_set_state", "fluid_solve_psi", "fluid_compute_velocity",
    "fluid_step", "fluid_max_speed",
    # particles.h
    "particles_init", "particles_free", "particles_color_from_image", "particles_advect",
    # render.h
    "render_init", "render_free", "render_fade", "render_deposit", "render_texel",
    "render_sample_dir", "render_pixel8",
    # noise.h
    "noise_init", "simplex3", "curl_noise",
    # ggb_ext.h
    "ggb_device_init", "ggb_version", "ggb_sync", "ggb_set_sync_mirrors", "fluid_sync_host", "fluid_sync_device",
    "fluid_get_f64", "fluid_set_f64", "fluid_solve_psi_device", "fluid_true_relres",
    "fluid_vcycle_f64", "ggb_laplacian_f64", "ggb_applyK_f64", "ggb_curl_f64",
    "ggb_divergence_f64", "ggb_stencil_compress", "particles_sync_host", "particles_get_f64", "particles_set_f64",
    "particles_set_round32", "render_sync_host", "render_last_entries", "render_last_stats", "render_faces8",
    "render_net8", "render_preview8",
    "fluid_adjoint_new", "fluid_adjoint_free", "fluid_adjoint_forward", "fluid_adjoint_backward",
    "fluid_adjoint_iterations", "fluid_adjoint_fix_iterations", "fluid_remove_mean_noalloc",
    # ggb_inputs.h
    "ggb_initial_velocity", "ggb_banded_strip", "ggb_rgba8_to_float",
]

_lib = None


def lib():
    """Load libggb.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libggb.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        fp, dp, ip = P(C.c_float), P(C.c_double), P(C.c_int)
        sig = {
            "cs_init": (C.c_int, [P(CCubeSphere), C.c_int]),
            "cs_free": (None, [P(CCubeSphere)]),
            "cs_cell_of": (C.c_int, [P(CCubeSphere), CVec3]),
            "cs_taps": (None, [P(CCubeSphere), CVec3, P(C.c_int), P(C.c_double)]),
            "fluid_init": (C.c_int, [P(CFluid), P(CCubeSphere), P(CFluidParams)]),
            "fluid_free": (None, [P(CFluid)]),
            "fluid_curl": (None, [P(CCubeSphere), fp, fp]),
            "fluid_divergence": (None, [P(CCubeSphere), fp, fp]),
            "fluid_laplacian": (None, [P(CCubeSphere), fp, fp]),
            "fluid_remove_mean": (None, [P(CCubeSphere), fp]),
            "fluid_set_state": (None, [P(CFluid), fp, fp]),
            "fluid_solve_psi": (C.c_int, [P(CFluid)]),
            "fluid_compute_velocity": (None, [P(CFluid)]),
            "fluid_step": (None, [P(CFluid), C.c_double]),
            "fluid_max_speed": (C.c_double, [P(CFluid)]),
            "fluid_sync_host": (None, [P(CFluid)]),
            "fluid_sync_device": (None, [P(CFluid)]),
            "fluid_get_f64": (C.c_int, [P(CFluid), C.c_int, dp]),
            "fluid_set_f64": (C.c_int, [P(CFluid), C.c_int, dp]),
            "fluid_solve_psi_device": (C.c_int, [P(CFluid)]),
            "fluid_true_relres": (C.c_double, [P(CFluid)]),
            "fluid_vcycle_f64": (C.c_int, [P(CFluid), dp, dp]),
            "ggb_laplacian_f64": (C.c_int, [P(CCubeSphere), dp, dp]),
            "ggb_applyK_f64": (C.c_int, [P(CCubeSphere), dp, dp]),
            "ggb_curl_f64": (C.c_int, [P(CCubeSphere), dp, dp]),
            "ggb_divergence_f64": (C.c_int, [P(CCubeSphere), dp, dp]),
            "ggb_stencil_compress": (C.c_int, [P(CCubeSphere), ip, fp, ip, ip, ip]),
            "particles_init": (C.c_int, [P(CParticles), C.c_int]),
            "particles_free": (None, [P(CParticles)]),
            "particles_color_from_image": (None, [P(CParticles), fp, C.c_int, C.c_int]),
            "particles_advect": (None, [P(CParticles), P(CCubeSphere), fp, C.c_double]),
            "particles_sync_host": (None, [P(CParticles)]),
            "particles_get_f64": (C.c_int, [P(CParticles), dp, dp]),
            "particles_set_f64": (C.c_int, [P(CParticles), dp, dp]),
            "particles_set_round32": (None, [P(CParticles), C.c_int]),
            "render_init": (C.c_int, [P(CRender), C.c_int, C.c_float, C.c_float]),
            "render_free": (None, [P(CRender)]),
            "render_fade": (None, [P(CRender), C.c_double]),
            "render_deposit": (None, [P(CRender), P(CParticles)]),
            "render_texel": (None, [P(CRender), C.c_int, C.c_int, C.c_int, fp]),
            "render_pixel8": (None, [fp, fp, P(C.c_ubyte)]),
            "render_sync_host": (None, [P(CRender)]),
            "render_last_entries": (C.c_longlong, [P(CRender)]),
            "render_faces8": (None, [P(CRender), fp, P(C.c_ubyte)]),
            "render_net8": (None, [P(CRender), fp, P(C.c_ubyte)]),
            "render_preview8": (C.c_int, [P(CRender), fp, C.c_int, C.c_int, P(C.c_ubyte)]),
            "render_sample_dir": (None, [P(CRender), C.c_double, C.c_double, C.c_double, fp]),
            "render_last_stats": (C.c_int, [P(CRender), P(C.c_longlong)]),
            "fluid_adjoint_new": (C.c_void_p, [P(CFluid), C.c_int]),
            "fluid_adjoint_free": (None, [C.c_void_p]),
            "fluid_adjoint_forward": (C.c_int, [C.c_void_p, dp, dp, C.c_int, dp]),
            "fluid_adjoint_backward": (C.c_int, [C.c_void_p, dp, dp, dp]),
            "fluid_adjoint_iterations": (C.c_int, [C.c_void_p, ip]),
            "fluid_adjoint_fix_iterations": (C.c_int, [C.c_void_p, ip, C.c_int]),
            "ggb_device_init": (C.c_int, []),
            "ggb_version": (C.c_char_p, []),
            "ggb_sync": (None, []),
            "ggb_set_sync_mirrors": (None, [C.c_int]),
            "ggb_initial_velocity": (C.c_int, [P(CCubeSphere), C.c_uint, C.c_int, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.c_int, fp, fp]),
            "ggb_banded_strip": (C.c_int, [C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, P(C.c_ubyte)]),
            "ggb_rgba8_to_float": (None, [P(C.c_ubyte), C.c_size_t, fp]),
            "ggbk_probe_enable": (None, [C.c_int, C.c_int]),
            "ggbk_probe_collect": (C.c_int, [C.c_int, P(C.c_double)]),
            "ggbk_copy_kernel": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
            "ggbk_malloc": (C.c_void_p, [C.c_size_t]),
            "ggbk_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
            "ggbk_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
            "fluid_remove_mean_noalloc": (C.c_int, [P(CCubeSphere), C.c_void_p]),
            "ggbk_free": (None, [C.c_void_p]),
            "ggbk_event_new": (C.c_void_p, []),
            "ggbk_event_record": (None, [C.c_void_p]),
            "ggbk_event_ms": (C.c_float, [C.c_void_p, C.c_void_p]),
            "ggbk_event_free": (None, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _view(ptr, count, dtype):
    if not ptr or count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)


class Grid:
    """CubeSphere built by the library's cs_init (host geometry + seam ghost table)."""

    def __init__(self, N):
        self.c = CCubeSphere()
        if lib().cs_init(C.byref(self.c), int(N)) != 0:
            raise RuntimeError("cs_init failed")
        n = self.c.ncells
        self.N, self.n, self.h = self.c.N, n, self.c.h
        self.total_area = self.c.total_area
        self.pos = _view(self.c.pos, 3 * n, np.float32).reshape(n, 3)
        self.area = _view(self.c.area, n, np.float32)
        self.nnb = _view(self.c.nnb, n, np.uint8)
        self.nbr = _view(self.c.nbr, CS_MAXNB * n, np.int32).reshape(n, CS_MAXNB)
        self.w = _view(self.c.w, CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB)
        self.elen = _view(self.c.elen, CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB)
        self.gc = _view(self.c.gc, 3 * CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB, 3)
        P = self.N + 2
        self.ghost = _view(self.c.ghost, 6 * P * P, np.int32).reshape(6, P, P)

    def free(self):
        if self.c.pos:
            lib().cs_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def cell_of(self, p):
        return lib().cs_cell_of(C.byref(self.c), CVec3(*map(float, p)))

    def taps(self, p):
        idx = (C.c_int * 4)()
        wt = (C.c_double * 4)()
        lib().cs_taps(C.byref(self.c), CVec3(*map(float, p)), idx, wt)
        return np.array(idx[:], np.int32), np.array(wt[:])

    def initial_velocity(self, seed=1, bands=8, band_amp=1.0, noise_amp=0.5, noise_scale=3.0,
                         lacunarity=2.0, gain=0.5, octaves=4):
        vel = np.zeros(3 * self.n, np.float32)
        vband = np.zeros(3 * self.n, np.float32)
        rc = lib().ggb_initial_velocity(C.byref(self.c), seed, bands, band_amp, noise_amp,
                                        noise_scale, lacunarity, gain, octaves, fptr(vel), fptr(vband))
        if rc:
            raise RuntimeError("initial_velocity failed")
        return vel, vband

    # free operators (host arrays; reference semantics, float)
    def laplacian(self, x):
        y = np.zeros(self.n, np.float32)
        lib().fluid_laplacian(C.byref(self.c), fptr(np.ascontiguousarray(x, np.float32)), fptr(y))
        return y

    def curl(self, vel):
        y = np.zeros(self.n, np.float32)
        lib().fluid_curl(C.byref(self.c), fptr(np.ascontiguousarray(vel, np.float32).ravel()), fptr(y))
        return y

    def divergence(self, vel):
        y = np.zeros(self.n, np.float32)
        lib().fluid_divergence(C.byref(self.c), fptr(np.ascontiguousarray(vel, np.float32).ravel()), fptr(y))
        return y

    def remove_mean(self, f):
        f = np.ascontiguousarray(f, np.float32).copy()
        lib().fluid_remove_mean(C.byref(self.c), fptr(f))
        return f

    def stencil_compressed(self):
        """(rec [n,2], class weights [ncls,8], explicit rows [nirr,8]) of the compressed
        stencil the device passes read."""
        n = self.n
        rec = np.zeros(2 * n, np.int32)
        cw = np.zeros(8 * n, np.float32)
        xn = np.zeros(8 * n, np.int32)
        ncls, nirr = C.c_int(), C.c_int()
        if lib().ggb_stencil_compress(C.byref(self.c), iptr(rec), fptr(cw), iptr(xn), C.byref(ncls),
                                      C.byref(nirr)):
            raise MemoryError("ggb_stencil_compress")
        return rec.reshape(n, 2), cw[:8 * ncls.value].reshape(-1, 8), xn[:8 * nirr.value].reshape(-1, 8)

    def laplacian_f64(self, x):
        y = np.zeros(self.n)
        lib().ggb_laplacian_f64(C.byref(self.c), dptr(np.ascontiguousarray(x, np.float64)), dptr(y))
        return y

    def applyK_f64(self, x):
        y = np.zeros(self.n)
        lib().ggb_applyK_f64(C.byref(self.c), dptr(np.ascontiguousarray(x, np.float64)), dptr(y))
        return y

    def curl_f64(self, vel):
        y = np.zeros(self.n)
        lib().ggb_curl_f64(C.byref(self.c), dptr(np.ascontiguousarray(vel, np.float64).ravel()), dptr(y))
        return y

    def divergence_f64(self, vel):
        y = np.zeros(self.n)
        lib().ggb_divergence_f64(C.byref(self.c), dptr(np.ascontiguousarray(vel, np.float64).ravel()), dptr(y))
        return y


class Fluid:
    def __init__(self, grid, forcing=0.0, rotation=0.0, bfecc=1, cg_tol=1e-4, cg_maxit=500,
                 multigrid=1, precision=PREC_FP64):
        self.grid = grid
        self.prm = CFluidParams(forcing, rotation, bfecc, cg_tol, cg_maxit, multigrid, precision)
        self.c = CFluid()
        if lib().fluid_init(C.byref(self.c), C.byref(grid.c), C.byref(self.prm)) != 0:
            raise RuntimeError("fluid_init failed")
        n = grid.n
        self.omega = _view(self.c.omega, n, np.float32)
        self.omega_band = _view(self.c.omega_band, n, np.float32)
        self.psi = _view(self.c.psi, n, np.float32)
        self.vel = _view(self.c.vel, 3 * n, np.float32)

    def free(self):
        if self.c.dev:
            lib().fluid_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def cg_iters(self):
        return self.c.cg_iters

    @property
    def cg_relres(self):
        return self.c.cg_relres

    @property
    def nlev(self):
        return self.c.nlev

    def set_state(self, vel0, vband=None):
        v0 = np.ascontiguousarray(vel0, np.float32).ravel()
        vb = None if vband is None else fptr(np.ascontiguousarray(vband, np.float32).ravel())
        lib().fluid_set_state(C.byref(self.c), fptr(v0), vb)

    def step(self, dt):
        lib().fluid_step(C.byref(self.c), float(dt))

    def max_speed(self):
        return lib().fluid_max_speed(C.byref(self.c))

    def solve_psi(self):
        return lib().fluid_solve_psi(C.byref(self.c))

    def solve_psi_device(self):
        return lib().fluid_solve_psi_device(C.byref(self.c))

    def compute_velocity(self):
        lib().fluid_compute_velocity(C.byref(self.c))

    def sync_host(self):
        lib().fluid_sync_host(C.byref(self.c))

    def get(self, field):
        n = self.grid.n * (3 if field in (FIELD_VEL, FIELD_DEP_B, FIELD_DEP_F) else 1)
        out = np.zeros(n)
        if lib().fluid_get_f64(C.byref(self.c), field, dptr(out)):
            raise ValueError(field)
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, np.float64).ravel()
        if lib().fluid_set_f64(C.byref(self.c), field, dptr(v)):
            raise ValueError(field)

    def true_relres(self):
        return lib().fluid_true_relres(C.byref(self.c))

    def set_forcing(self, k):
        """Forcing strength used by later steps (F->prm.forcing)."""
        self.c.prm.forcing = float(k)

    def vcycle(self, b):
        x = np.zeros(self.grid.n)
        if lib().fluid_vcycle_f64(C.byref(self.c), dptr(np.ascontiguousarray(b, np.float64)), dptr(x)):
            raise RuntimeError("vcycle failed")
        return x


class FluidAdjoint:
    """Checkpointed discrete adjoint of a rollout of fluid_step (fluid_adjoint_* ABI)."""

    def __init__(self, fluid, checkpoint_every=10):
        self.fluid = fluid
        self.h = lib().fluid_adjoint_new(C.byref(fluid.c), int(checkpoint_every))
        if not self.h:
            raise RuntimeError("fluid_adjoint_new failed (needs FP64 state)")

    def free(self):
        if self.h:
            lib().fluid_adjoint_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def forward(self, omega0, dts):
        """omega_T after len(dts) steps from the control omega0."""
        om = np.ascontiguousarray(omega0, np.float64).ravel()
        d = np.ascontiguousarray(dts, np.float64).ravel()
        out = np.zeros(self.fluid.grid.n)
        if lib().fluid_adjoint_forward(self.h, dptr(om), dptr(d), len(d), dptr(out)) != 0:
            raise RuntimeError("fluid_adjoint_forward failed")
        return out

    def backward(self, g_omega_T):
        """(dL/d omega0, dL/d forcing) from dL/d omega_T, for the last forward."""
        g = np.ascontiguousarray(g_omega_T, np.float64).ravel()
        g0 = np.zeros(self.fluid.grid.n)
        gk = C.c_double()
        if lib().fluid_adjoint_backward(self.h, dptr(g), dptr(g0), C.byref(gk)) != 0:
            raise RuntimeError("fluid_adjoint_backward failed")
        return g0, gk.value

    def iterations(self):
        """CG iteration counts of the last forward: [set-up solve, step 0, step 1, ...]."""
        n = lib().fluid_adjoint_iterations(self.h, None)
        if n < 0:
            raise RuntimeError("no forward pass yet")
        its = np.zeros(n, np.int32)
        lib().fluid_adjoint_iterations(self.h, iptr(its))
        return its

    def fix_iterations(self, its):
        """Prescribe every solve's iteration count for later forwards (None: tolerance)."""
        if its is None:
            rc = lib().fluid_adjoint_fix_iterations(self.h, None, 0)
        else:
            a = np.ascontiguousarray(its, np.int32)
            rc = lib().fluid_adjoint_fix_iterations(self.h, iptr(a), len(a))
        if rc != 0:
            raise ValueError("bad iteration schedule")


class Particles:
    def __init__(self, n):
        self.c = CParticles()
        if lib().particles_init(C.byref(self.c), int(n)) != 0:
            raise RuntimeError("particles_init failed")
        self.n = n
        self.pos = _view(self.c.pos, 3 * n, np.float32)
        self.prev = _view(self.c.prev, 3 * n, np.float32)
        self.col = _view(self.c.col, 4 * n, np.float32)

    def free(self):
        if self.c.dev:
            lib().particles_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def color_from_image(self, img_f32, w, h):
        lib().particles_color_from_image(C.byref(self.c), fptr(np.ascontiguousarray(img_f32, np.float32).ravel()), w, h)

    def advect(self, grid, vel, dt):
        """vel: a Fluid (device velocity) or a float32 host array"""
        if isinstance(vel, Fluid):
            lib().particles_advect(C.byref(self.c), C.byref(grid.c), vel.c.vel, float(dt))
        else:
            v = np.ascontiguousarray(vel, np.float32).ravel()
            lib().particles_advect(C.byref(self.c), C.byref(grid.c), fptr(v), float(dt))

    def get(self):
        pos = np.zeros(3 * self.n)
        prev = np.zeros(3 * self.n)
        lib().particles_get_f64(C.byref(self.c), dptr(pos), dptr(prev))
        return pos, prev

    def set(self, pos, prev):
        lib().particles_set_f64(C.byref(self.c), dptr(np.ascontiguousarray(pos, np.float64).ravel()),
                                dptr(np.ascontiguousarray(prev, np.float64).ravel()))

    def set_round32(self, on):
        lib().particles_set_round32(C.byref(self.c), int(bool(on)))

    def sync_host(self):
        lib().particles_sync_host(C.byref(self.c))


class Render:
    def __init__(self, R, radius=1.0, alpha=0.35):
        self.c = CRender()
        if lib().render_init(C.byref(self.c), int(R), float(radius), float(alpha)) != 0:
            raise RuntimeError("render_init failed")
        self.R = R
        self.buf = _view(self.c.buf, 6 * R * R * 4, np.float32).reshape(6, R, R, 4)

    def free(self):
        if self.c.dev:
            lib().render_free(C.byref(self.c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def g(self):
        return self.c.g

    def fade(self, f):
        lib().render_fade(C.byref(self.c), float(f))

    def deposit(self, particles):
        lib().render_deposit(C.byref(self.c), C.byref(particles.c))

    def sync_host(self):
        lib().render_sync_host(C.byref(self.c))
        return self.buf

    def last_entries(self):
        return lib().render_last_entries(C.byref(self.c))

    def last_stats(self):
        out = (C.c_longlong * 16)()
        lib().render_last_stats(C.byref(self.c), out)
        return list(out)

    def faces8(self, bg=None):
        ch = 4 if bg is None else 3
        out = np.zeros((6, self.R, self.R, ch), np.uint8)
        bgp = None if bg is None else fptr(np.asarray(bg, np.float32))
        lib().render_faces8(C.byref(self.c), bgp, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out

    def net8(self, bg=None):
        out = np.zeros((3 * self.R, 4 * self.R, 4), np.uint8)
        bgp = None if bg is None else fptr(np.asarray(bg, np.float32))
        lib().render_net8(C.byref(self.c), bgp, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out

    def preview8(self, W=None, H=None, bg=None):
        W = 2 * self.R if W is None else W
        H = self.R if H is None else H
        out = np.zeros((H, W, 4 if bg is None else 3), np.uint8)
        bgp = None if bg is None else fptr(np.asarray(bg, np.float32))
        if lib().render_preview8(C.byref(self.c), bgp, W, H, out.ctypes.data_as(C.POINTER(C.c_ubyte))):
            raise ValueError("bad preview size")
        return out

    def sample_dir(self, d):
        """render_sample_dir (host reader over the lazily refreshed mirror)."""
        out = np.zeros(4, np.float32)
        lib().render_sample_dir(C.byref(self.c), float(d[0]), float(d[1]), float(d[2]), fptr(out))
        return out

    def texel(self, f, x, y):
        out = np.zeros(4, np.float32)
        lib().render_texel(C.byref(self.c), int(f), int(x), int(y), fptr(out))
        return out


def banded_strip(seed, w=64, h=1024, nbands=24, hmin=16, hmax=64):
    img = np.zeros((h, w, 4), np.uint8)
    rc = lib().ggb_banded_strip(seed, w, h, nbands, hmin, hmax, img.ctypes.data_as(C.POINTER(C.c_ubyte)))
    if rc:
        raise ValueError("bad strip parameters")
    return img


def strip_to_float(img8):
    return (img8.astype(np.float32) / np.float32(255.0)).ravel()


def device_init():
    return lib().ggb_device_init()


class Timer:
    """HIP events on the library stream (the stream every kernel is launched on)."""

    def __init__(self):
        L = lib()
        self.a = L.ggbk_event_new()
        self.b = L.ggbk_event_new()

    def start(self):
        lib().ggbk_event_record(self.a)

    def stop_ms(self):
        lib().ggbk_event_record(self.b)
        return lib().ggbk_event_ms(self.a, self.b)

    def __del__(self):
        try:
            lib().ggbk_event_free(self.a)
            lib().ggbk_event_free(self.b)
        except Exception:
            pass
