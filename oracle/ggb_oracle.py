"""CPU oracle bindings -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's CPU
baseline).  Never imported by the product.

Two backends:
  * the C restatement (oracle/ggb_oracle.c), built twice: "f64" (double storage, the
    FP64 oracle) and "f32" (float storage, the reference's precision);
  * the reference itself ("ref"): its numerical core compiled from its own sources by
    oracle/Makefile into oracle/_ref/libggbref.so (parity flags) and
    libggbref_fast.so (optimised, the CPU timing baseline).
"""
import ctypes as C
import ctypes.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "f64": os.path.join(HERE, "lib", "libggb_oracle_f64.so"),
    "f32": os.path.join(HERE, "lib", "libggb_oracle_f32.so"),
    "ref": os.path.join(HERE, "_ref", "libggbref.so"),
    "ref_fast": os.path.join(HERE, "_ref", "libggbref_fast.so"),
}
CS_MAXNB = 8


def available(kind):
    return os.path.exists(LIBS[kind])


def set_omp_threads(n):
    """Thread count of the process's OpenMP runtime (shared by every loaded library)."""
    name = ctypes.util.find_library("gomp") or "libgomp.so.1"
    C.CDLL(name).omp_set_num_threads(int(n))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- restatement

class OGrid(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ncells", C.c_int), ("h", C.c_double), ("total_area", C.c_double),
        ("pos", C.POINTER(C.c_float)), ("area", C.POINTER(C.c_float)),
        ("nnb", C.POINTER(C.c_ubyte)), ("nbr", C.POINTER(C.c_int)),
        ("w", C.POINTER(C.c_float)), ("elen", C.POINTER(C.c_float)),
        ("gc", C.POINTER(C.c_float)),
    ]


class ORender(C.Structure):
    _fields_ = [("R", C.c_int), ("buf", C.POINTER(C.c_float)), ("g", C.c_double),
                ("radius", C.c_float), ("alpha", C.c_float)]


_olibs = {}


def olib(kind):
    if kind not in _olibs:
        path = LIBS[kind]
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: %s (run make -C oracle)" % path)
        L = C.CDLL(path)
        P = C.POINTER
        real = C.c_double if kind == "f64" else C.c_float
        rp = P(real)
        fp, dp, ip = P(C.c_float), P(C.c_double), P(C.c_int)
        g = P(OGrid)
        sig = {
            "orc_mg_nlev": (C.c_int, [C.c_int]),
            "orc_taps": (None, [g, dp, ip, dp]),
            "orc_cell_of": (C.c_int, [g, dp]),
            "orc_fluid_new": (C.c_void_p, [g, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int]),
            "orc_fluid_free": (None, [C.c_void_p]),
            "orc_fluid_nlev": (C.c_int, [C.c_void_p]),
            "orc_fluid_cg_iters": (C.c_int, [C.c_void_p]),
            "orc_fluid_cg_relres": (C.c_double, [C.c_void_p]),
            "orc_fluid_transfer": (None, [C.c_void_p, C.c_int, ip, fp]),
            "orc_divergence": (None, [g, rp, rp]),
            "orc_curl": (None, [g, rp, rp]),
            "orc_laplacian": (None, [g, rp, rp]),
            "orc_applyK": (None, [g, rp, dp]),
            "orc_remove_mean": (None, [g, rp]),
            "orc_solve_psi": (C.c_int, [C.c_void_p]),
            "orc_compute_velocity": (None, [C.c_void_p]),
            "orc_fluid_set_state": (None, [C.c_void_p, fp, fp]),
            "orc_fluid_step": (None, [C.c_void_p, C.c_double]),
            "orc_fluid_max_speed": (C.c_double, [C.c_void_p]),
            "orc_fluid_get": (C.c_int, [C.c_void_p, C.c_int, dp]),
            "orc_fluid_set": (C.c_int, [C.c_void_p, C.c_int, dp]),
            "orc_vcycle": (None, [C.c_void_p, rp, rp]),
            "orc_fluid_set_forcing": (None, [C.c_void_p, C.c_double]),
            "orc_adj_forward": (C.c_int, [C.c_void_p, dp, dp, C.c_int]),
            "orc_adj_backward": (C.c_int, [C.c_void_p, dp, dp, dp]),
            "orc_adj_release": (None, [C.c_void_p]),
            "orc_particles_init": (None, [C.c_int, rp, rp, fp]),
            "orc_particles_color": (None, [C.c_int, rp, fp, fp, C.c_int, C.c_int]),
            "orc_particles_advect": (None, [g, C.c_int, rp, rp, rp, C.c_double]),
            "orc_render_init": (None, [P(ORender), C.c_int, C.c_float, C.c_float, fp]),
            "orc_render_fade": (None, [P(ORender), C.c_double]),
            "orc_render_deposit": (None, [P(ORender), C.c_int, rp, rp, fp]),
            "orc_render_count_records": (C.c_longlong, [P(ORender), C.c_int, rp, rp]),
            "orc_render_faces8": (None, [P(ORender), fp, P(C.c_ubyte)]),
            "orc_render_texel": (None, [P(ORender), C.c_int, C.c_int, C.c_int, fp]),
            "orc_render_pixel8": (None, [fp, fp, P(C.c_ubyte)]),
            "orc_render_sample_dir": (None, [P(ORender), C.c_double, C.c_double, C.c_double, fp]),
            "orc_render_net8": (None, [P(ORender), fp, P(C.c_ubyte)]),
            "orc_render_preview8": (None, [P(ORender), C.c_int, C.c_int, fp, P(C.c_ubyte)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.real = real
        L.np_real = np.float64 if kind == "f64" else np.float32
        _olibs[kind] = L
    return _olibs[kind]


def ogrid_from(grid):
    """OGrid view of a grid object exposing numpy arrays (ggb.Grid or RefGrid)."""
    og = OGrid()
    og.N, og.ncells, og.h, og.total_area = grid.N, grid.n, grid.h, grid.total_area
    og.pos = _fp(grid.pos)
    og.area = _fp(grid.area)
    og.nnb = grid.nnb.ctypes.data_as(C.POINTER(C.c_ubyte))
    og.nbr = grid.nbr.ctypes.data_as(C.POINTER(C.c_int))
    og.w = _fp(grid.w)
    og.elen = _fp(grid.elen)
    og.gc = _fp(grid.gc)
    og._keep = grid
    return og


class Oracle:
    """One precision of the restatement ("f64" or "f32")."""

    def __init__(self, kind="f64"):
        self.kind = kind
        self.L = olib(kind)
        self.dt = self.L.np_real

    def arr(self, a):
        return np.ascontiguousarray(a, self.dt)

    def rp(self, a):
        return a.ctypes.data_as(C.POINTER(self.L.real))

    def laplacian(self, grid, x):
        og = ogrid_from(grid)
        x = self.arr(x)
        y = np.zeros(grid.n, self.dt)
        self.L.orc_laplacian(C.byref(og), self.rp(x), self.rp(y))
        return y

    def applyK(self, grid, x):
        og = ogrid_from(grid)
        x = self.arr(x)
        y = np.zeros(grid.n)
        self.L.orc_applyK(C.byref(og), self.rp(x), _dp(y))
        return y

    def curl(self, grid, vel):
        og = ogrid_from(grid)
        v = self.arr(vel).ravel()
        y = np.zeros(grid.n, self.dt)
        self.L.orc_curl(C.byref(og), self.rp(v), self.rp(y))
        return y

    def divergence(self, grid, vel):
        og = ogrid_from(grid)
        v = self.arr(vel).ravel()
        y = np.zeros(grid.n, self.dt)
        self.L.orc_divergence(C.byref(og), self.rp(v), self.rp(y))
        return y

    def remove_mean(self, grid, f):
        og = ogrid_from(grid)
        f = self.arr(f).copy()
        self.L.orc_remove_mean(C.byref(og), self.rp(f))
        return f

    def cell_of(self, grid, p):
        og = ogrid_from(grid)
        return self.L.orc_cell_of(C.byref(og), _dp(np.ascontiguousarray(p, np.float64)))

    def taps(self, grid, p):
        og = ogrid_from(grid)
        idx = np.zeros(4, np.int32)
        wt = np.zeros(4)
        self.L.orc_taps(C.byref(og), _dp(np.ascontiguousarray(p, np.float64)),
                        idx.ctypes.data_as(C.POINTER(C.c_int)), _dp(wt))
        return idx, wt


class OracleFluid:
    """Restated Fluid on a stack of grids (levels[0] fine; coarse grids N >> l)."""

    def __init__(self, kind, levels, forcing=0.0, rotation=0.0, bfecc=1, cg_tol=1e-4, cg_maxit=500,
                 multigrid=1):
        self.o = Oracle(kind)
        self.levels = levels
        self.og = (OGrid * len(levels))(*[ogrid_from(g) for g in levels])
        nlev = self.o.L.orc_mg_nlev(levels[0].N)
        if multigrid and nlev > len(levels):
            raise ValueError("need %d grid levels" % nlev)
        self.h = self.o.L.orc_fluid_new(self.og, forcing, rotation, bfecc, cg_tol, cg_maxit, multigrid)
        self.n = levels[0].n

    def __del__(self):
        try:
            self.o.L.orc_adj_release(self.h)
            self.o.L.orc_fluid_free(self.h)
        except Exception:
            pass

    def set_forcing(self, k):
        self.o.L.orc_fluid_set_forcing(self.h, float(k))

    def adj_forward(self, omega0, dts):
        """Rollout from the control omega0 with the dt schedule; returns omega_T."""
        om = np.ascontiguousarray(omega0, np.float64).ravel()
        d = np.ascontiguousarray(dts, np.float64).ravel()
        if self.o.L.orc_adj_forward(self.h, _dp(om), _dp(d), len(d)) != 0:
            raise RuntimeError("orc_adj_forward failed")
        return self.get(0)

    def adj_backward(self, gT):
        """dL/d omega0 and dL/d forcing from dL/d omega_T (after adj_forward)."""
        g = np.ascontiguousarray(gT, np.float64).ravel()
        out = np.zeros(self.n)
        gk = C.c_double()
        if self.o.L.orc_adj_backward(self.h, _dp(g), _dp(out), C.byref(gk)) != 0:
            raise RuntimeError("orc_adj_backward failed")
        return out, gk.value

    def set_state(self, vel0, vband=None):
        v0 = np.ascontiguousarray(vel0, np.float32).ravel()
        vb = None if vband is None else _fp(np.ascontiguousarray(vband, np.float32).ravel())
        self.o.L.orc_fluid_set_state(self.h, _fp(v0), vb)

    def step(self, dt):
        self.o.L.orc_fluid_step(self.h, float(dt))

    def solve_psi(self):
        return self.o.L.orc_solve_psi(self.h)

    def compute_velocity(self):
        self.o.L.orc_compute_velocity(self.h)

    def max_speed(self):
        return self.o.L.orc_fluid_max_speed(self.h)

    @property
    def cg_iters(self):
        return self.o.L.orc_fluid_cg_iters(self.h)

    @property
    def cg_relres(self):
        return self.o.L.orc_fluid_cg_relres(self.h)

    def get(self, field):
        m = 3 if field in (3, 4, 5) else 1
        out = np.zeros(m * self.n)
        self.o.L.orc_fluid_get(self.h, field, _dp(out))
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, np.float64).ravel()
        self.o.L.orc_fluid_set(self.h, field, _dp(v))

    def vcycle(self, b):
        b = self.o.arr(b)
        x = np.zeros(self.n, self.o.dt)
        self.o.L.orc_vcycle(self.h, self.o.rp(b), self.o.rp(x))
        return x

    def transfer(self, l):
        n = self.levels[l].n
        ptap = np.zeros(4 * n, np.int32)
        pw = np.zeros(4 * n, np.float32)
        self.o.L.orc_fluid_transfer(self.h, l, ptap.ctypes.data_as(C.POINTER(C.c_int)), _fp(pw))
        return ptap.reshape(n, 4), pw.reshape(n, 4)


class OracleParticles:
    def __init__(self, kind, n):
        self.o = Oracle(kind)
        self.n = n
        self.pos = np.zeros(3 * n, self.o.dt)
        self.prev = np.zeros(3 * n, self.o.dt)
        self.col = np.zeros(4 * n, np.float32)
        self.o.L.orc_particles_init(n, self.o.rp(self.pos), self.o.rp(self.prev), _fp(self.col))

    def color_from_image(self, img_f32, w, h):
        self.o.L.orc_particles_color(self.n, self.o.rp(self.pos), _fp(self.col),
                                     _fp(np.ascontiguousarray(img_f32, np.float32).ravel()), w, h)

    def advect(self, grid, vel, dt):
        og = ogrid_from(grid)
        v = self.o.arr(vel).ravel()
        self.o.L.orc_particles_advect(C.byref(og), self.n, self.o.rp(self.pos), self.o.rp(self.prev),
                                      self.o.rp(v), float(dt))


class OracleRender:
    def __init__(self, kind, R, radius=1.0, alpha=0.35):
        self.o = Oracle(kind)
        self.R = R
        self.buf = np.zeros((6, R, R, 4), np.float32)
        self.r = ORender()
        self.o.L.orc_render_init(C.byref(self.r), R, radius, alpha, _fp(self.buf))

    @property
    def g(self):
        return self.r.g

    def fade(self, f):
        self.o.L.orc_render_fade(C.byref(self.r), float(f))

    def deposit(self, parts):
        self.o.L.orc_render_deposit(C.byref(self.r), parts.n, self.o.rp(parts.prev), self.o.rp(parts.pos),
                                    _fp(parts.col))

    def count_records(self, parts):
        return self.o.L.orc_render_count_records(C.byref(self.r), parts.n, self.o.rp(parts.prev),
                                                 self.o.rp(parts.pos))

    def _bg(self, bg):
        if bg is None:
            return None, 4
        self._bgkeep = np.ascontiguousarray(bg, np.float32)
        return _fp(self._bgkeep), 3

    def faces8(self, bg=None):
        bgp, ch = self._bg(bg)
        out = np.zeros((6, self.R, self.R, ch), np.uint8)
        self.o.L.orc_render_faces8(C.byref(self.r), bgp, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out

    def sample_dir(self, d):
        out = np.zeros(4, np.float32)
        self.o.L.orc_render_sample_dir(C.byref(self.r), float(d[0]), float(d[1]), float(d[2]), _fp(out))
        return out

    def net8(self, bg=None):
        bgp, _ = self._bg(bg)
        out = np.zeros((3 * self.R, 4 * self.R, 4), np.uint8)
        self.o.L.orc_render_net8(C.byref(self.r), bgp, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out

    def preview8(self, W=None, H=None, bg=None):
        W = 2 * self.R if W is None else W
        H = self.R if H is None else H
        bgp, ch = self._bg(bg)
        out = np.zeros((H, W, ch), np.uint8)
        self.o.L.orc_render_preview8(C.byref(self.r), W, H, bgp, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out


# ---------------------------------------------------------------- compiled reference

class RCubeSphere(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ncells", C.c_int), ("h", C.c_double),
        ("pos", C.POINTER(C.c_float)), ("area", C.POINTER(C.c_float)),
        ("nnb", C.POINTER(C.c_ubyte)), ("nbr", C.POINTER(C.c_int)),
        ("w", C.POINTER(C.c_float)), ("elen", C.POINTER(C.c_float)),
        ("gc", C.POINTER(C.c_float)), ("total_area", C.c_double),
        ("asym_count", C.c_int), ("open_cells", C.c_int),
    ]


class RFluidParams(C.Structure):
    _fields_ = [("forcing", C.c_double), ("rotation", C.c_double), ("bfecc", C.c_int),
                ("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("multigrid", C.c_int)]


class RMGLevel(C.Structure):
    _fields_ = [("cs", RCubeSphere)] + [(k, C.c_void_p) for k in
                                        ("x", "b", "r", "dinv", "ptap", "pw", "rstart", "ridx", "rw")]


class RFluid(C.Structure):
    _fields_ = [("cs", C.POINTER(RCubeSphere)), ("prm", RFluidParams), ("nlev", C.c_int),
                ("lev", C.POINTER(RMGLevel))] + \
               [(k, C.POINTER(C.c_float)) for k in ("omega", "omega_band", "psi", "vel", "q0", "q1", "q2",
                                                    "dep_b", "dep_f", "r", "p", "Ap", "z", "dinv")] + \
               [("cg_iters", C.c_int), ("cg_relres", C.c_double)]


class RParticles(C.Structure):
    _fields_ = [("n", C.c_int), ("pos", C.POINTER(C.c_float)), ("prev", C.POINTER(C.c_float)),
                ("col", C.POINTER(C.c_float))]


class RRender(C.Structure):
    _fields_ = [("R", C.c_int), ("buf", C.POINTER(C.c_float)), ("g", C.c_double),
                ("radius", C.c_float), ("alpha", C.c_float), ("strip_h", C.c_int),
                ("nstrips", C.c_int), ("nbins", C.c_int), ("maxthreads", C.c_int),
                ("tr", C.c_void_p), ("sorted", C.c_void_p), ("sorted_cap", C.c_size_t),
                ("bin_start", C.c_void_p)]


class RNoise(C.Structure):
    _fields_ = [("perm", C.c_ubyte * 512)]


class RVec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class RCurlParams(C.Structure):
    _fields_ = [("scale", C.c_double), ("lacunarity", C.c_double), ("gain", C.c_double),
                ("octaves", C.c_int)]


_rlibs = {}


def rlib(kind="ref"):
    if kind not in _rlibs:
        path = LIBS[kind]
        if not os.path.exists(path):
            raise RuntimeError("reference build missing: %s" % path)
        L = C.CDLL(path)
        P = C.POINTER
        fp = P(C.c_float)
        sig = {
            "cs_init": (C.c_int, [P(RCubeSphere), C.c_int]),
            "cs_free": (None, [P(RCubeSphere)]),
            "cs_taps": (None, [P(RCubeSphere), RVec3, P(C.c_int), P(C.c_double)]),
            "cs_cell_of": (C.c_int, [P(RCubeSphere), RVec3]),
            "fluid_init": (C.c_int, [P(RFluid), P(RCubeSphere), P(RFluidParams)]),
            "fluid_free": (None, [P(RFluid)]),
            "fluid_curl": (None, [P(RCubeSphere), fp, fp]),
            "fluid_divergence": (None, [P(RCubeSphere), fp, fp]),
            "fluid_laplacian": (None, [P(RCubeSphere), fp, fp]),
            "fluid_remove_mean": (None, [P(RCubeSphere), fp]),
            "fluid_set_state": (None, [P(RFluid), fp, fp]),
            "fluid_solve_psi": (C.c_int, [P(RFluid)]),
            "fluid_compute_velocity": (None, [P(RFluid)]),
            "fluid_step": (None, [P(RFluid), C.c_double]),
            "fluid_max_speed": (C.c_double, [P(RFluid)]),
            "particles_init": (C.c_int, [P(RParticles), C.c_int]),
            "particles_free": (None, [P(RParticles)]),
            "particles_color_from_image": (None, [P(RParticles), fp, C.c_int, C.c_int]),
            "particles_advect": (None, [P(RParticles), P(RCubeSphere), fp, C.c_double]),
            "render_init": (C.c_int, [P(RRender), C.c_int, C.c_float, C.c_float]),
            "render_free": (None, [P(RRender)]),
            "render_fade": (None, [P(RRender), C.c_double]),
            "render_deposit": (None, [P(RRender), P(RParticles)]),
            "render_texel": (None, [P(RRender), C.c_int, C.c_int, C.c_int, fp]),
            "render_sample_dir": (None, [P(RRender), C.c_double, C.c_double, C.c_double, fp]),
            "render_pixel8": (None, [fp, fp, P(C.c_ubyte)]),
            "noise_init": (None, [P(RNoise), C.c_uint]),
            "simplex3": (C.c_double, [P(RNoise), RVec3, P(RVec3)]),
            "curl_noise": (RVec3, [P(RNoise), RVec3, P(RCurlParams)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _rlibs[kind] = L
    return _rlibs[kind]


def _rview(ptr, count, dtype):
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)


class RefGrid:
    """The reference's own cs_init (compiled from its sources)."""

    def __init__(self, N, kind="ref"):
        self.L = rlib(kind)
        self.c = RCubeSphere()
        if self.L.cs_init(C.byref(self.c), int(N)):
            raise RuntimeError("reference cs_init failed")
        n = self.c.ncells
        self.N, self.n, self.h, self.total_area = self.c.N, n, self.c.h, self.c.total_area
        self.asym_count, self.open_cells = self.c.asym_count, self.c.open_cells
        self.pos = _rview(self.c.pos, 3 * n, np.float32).reshape(n, 3)
        self.area = _rview(self.c.area, n, np.float32)
        self.nnb = _rview(self.c.nnb, n, np.uint8)
        self.nbr = _rview(self.c.nbr, CS_MAXNB * n, np.int32).reshape(n, CS_MAXNB)
        self.w = _rview(self.c.w, CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB)
        self.elen = _rview(self.c.elen, CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB)
        self.gc = _rview(self.c.gc, 3 * CS_MAXNB * n, np.float32).reshape(n, CS_MAXNB, 3)

    def __del__(self):
        try:
            self.L.cs_free(C.byref(self.c))
        except Exception:
            pass

    def taps(self, p):
        idx = (C.c_int * 4)()
        wt = (C.c_double * 4)()
        self.L.cs_taps(C.byref(self.c), RVec3(*map(float, p)), idx, wt)
        return np.array(idx[:], np.int32), np.array(wt[:])


class RefFluid:
    def __init__(self, grid, forcing=0.0, rotation=0.0, bfecc=1, cg_tol=1e-4, cg_maxit=500, multigrid=1):
        self.grid = grid
        self.L = grid.L
        self.prm = RFluidParams(forcing, rotation, bfecc, cg_tol, cg_maxit, multigrid)
        self.c = RFluid()
        self.L.fluid_init(C.byref(self.c), C.byref(grid.c), C.byref(self.prm))
        n = grid.n
        self.omega = _rview(self.c.omega, n, np.float32)
        self.omega_band = _rview(self.c.omega_band, n, np.float32)
        self.psi = _rview(self.c.psi, n, np.float32)
        self.vel = _rview(self.c.vel, 3 * n, np.float32)
        self.dep_b = _rview(self.c.dep_b, 3 * n, np.float32)

    def __del__(self):
        try:
            self.L.fluid_free(C.byref(self.c))
        except Exception:
            pass

    def set_state(self, vel0, vband=None):
        v0 = np.ascontiguousarray(vel0, np.float32).ravel()
        vb = None if vband is None else _fp(np.ascontiguousarray(vband, np.float32).ravel())
        self.L.fluid_set_state(C.byref(self.c), _fp(v0), vb)

    def step(self, dt):
        self.L.fluid_step(C.byref(self.c), float(dt))

    def solve_psi(self):
        return self.L.fluid_solve_psi(C.byref(self.c))

    def max_speed(self):
        return self.L.fluid_max_speed(C.byref(self.c))

    @property
    def cg_iters(self):
        return self.c.cg_iters

    @property
    def cg_relres(self):
        return self.c.cg_relres

    def transfer(self, l):
        lev = self.c.lev[l]
        n = lev.cs.ncells if l > 0 else self.grid.n
        ptap = np.ctypeslib.as_array(C.cast(lev.ptap, C.POINTER(C.c_int)), shape=(4 * n,)).copy()
        pw = np.ctypeslib.as_array(C.cast(lev.pw, C.POINTER(C.c_float)), shape=(4 * n,)).copy()
        return ptap.reshape(n, 4), pw.reshape(n, 4)


class RefParticles:
    def __init__(self, n, kind="ref"):
        self.L = rlib(kind)
        self.c = RParticles()
        if self.L.particles_init(C.byref(self.c), int(n)):
            raise RuntimeError("reference particles_init failed")
        self.n = n
        self.pos = _rview(self.c.pos, 3 * n, np.float32)
        self.prev = _rview(self.c.prev, 3 * n, np.float32)
        self.col = _rview(self.c.col, 4 * n, np.float32)

    def __del__(self):
        try:
            self.L.particles_free(C.byref(self.c))
        except Exception:
            pass

    def color_from_image(self, img_f32, w, h):
        self.L.particles_color_from_image(C.byref(self.c), _fp(np.ascontiguousarray(img_f32, np.float32).ravel()),
                                          w, h)

    def advect(self, grid, vel, dt):
        self.L.particles_advect(C.byref(self.c), C.byref(grid.c), _fp(np.ascontiguousarray(vel, np.float32)),
                                float(dt))


class RefRender:
    def __init__(self, R, radius=1.0, alpha=0.35, kind="ref"):
        self.L = rlib(kind)
        self.c = RRender()
        if self.L.render_init(C.byref(self.c), int(R), float(radius), float(alpha)):
            raise RuntimeError("reference render_init failed")
        self.R = R
        self.buf = _rview(self.c.buf, 6 * R * R * 4, np.float32).reshape(6, R, R, 4)

    def __del__(self):
        try:
            self.L.render_free(C.byref(self.c))
        except Exception:
            pass

    @property
    def g(self):
        return self.c.g

    def fade(self, f):
        self.L.render_fade(C.byref(self.c), float(f))

    def deposit(self, parts):
        self.L.render_deposit(C.byref(self.c), C.byref(parts.c))

    def texel(self, f, x, y):
        out = np.zeros(4, np.float32)
        self.L.render_texel(C.byref(self.c), int(f), int(x), int(y), _fp(out))
        return out

    def sample_dir(self, d):
        out = np.zeros(4, np.float32)
        self.L.render_sample_dir(C.byref(self.c), float(d[0]), float(d[1]), float(d[2]), _fp(out))
        return out

    def pixel8(self, c, bg=None):
        """render_pixel8 of one float texel: 4 bytes without bg, 3 with (the reference
        writes only those)."""
        c = np.ascontiguousarray(c, np.float32)
        out = np.zeros(4, np.uint8)
        bgk = None if bg is None else np.ascontiguousarray(bg, np.float32)
        self.L.render_pixel8(_fp(c), None if bgk is None else _fp(bgk), out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out if bg is None else out[:3]


class RefNoise:
    def __init__(self, seed, kind="ref"):
        self.L = rlib(kind)
        self.nz = RNoise()
        self.L.noise_init(C.byref(self.nz), int(seed))

    @property
    def perm(self):
        return np.array(self.nz.perm[:], np.uint8)

    def simplex3(self, p):
        g = RVec3()
        v = self.L.simplex3(C.byref(self.nz), RVec3(*map(float, p)), C.byref(g))
        return v, np.array([g.x, g.y, g.z])

    def curl(self, p, scale=3.0, lac=2.0, gain=0.5, octaves=4):
        cp = RCurlParams(scale, lac, gain, octaves)
        r = self.L.curl_noise(C.byref(self.nz), RVec3(*map(float, p)), C.byref(cp))
        return np.array([r.x, r.y, r.z])
