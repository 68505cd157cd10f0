"""ctypes bindings of libbmp_pic.so — the stage API of the gfx950 PIC step.

The shared library is loaded from this package's lib/ directory; if it is
missing, importing raises: there is no CPU fallback.

`Context` mirrors the reference's per-stage entry points (bin, deposit,
Poisson, gather + push, geometry) over numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libbmp_pic.so"

PIC_OK = 0
PIC_ERR_STATE = -13
PIC_LAYOUT_F64 = 0
PIC_LAYOUT_CELL32 = 1
GRID_RHO, GRID_PHI, GRID_EX, GRID_EY, GRID_BZ = 0, 1, 2, 3, 4
GRID_FINAL0, GRID_ENERGY0 = 8, 16
STAGES = ("deposit", "poisson", "efield", "push", "bin", "scatter", "finalize")

ERRORS = {
    -1: "PIC_ERR_ARG", -2: "PIC_ERR_FILE", -3: "PIC_ERR_FORMAT", -4: "PIC_ERR_ALLOC", -5: "PIC_ERR_SOLVER",
    -9: "PIC_ERR_PERIODIC_GEOM", -10: "PIC_ERR_HIP", -11: "PIC_ERR_RANGE", -12: "PIC_ERR_UNSUPPORTED",
    -13: "PIC_ERR_STATE",
}


class PicError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int),
        ("width", C.c_double), ("height", C.c_double),
        ("dt", C.c_double), ("epsilon", C.c_double),
        ("b0", C.c_double), ("bx", C.c_double), ("by", C.c_double),
        ("surf_area", C.c_double),
        ("periodic", C.c_int), ("layout", C.c_int),
        ("charge_shift", C.c_int), ("energy_shift", C.c_int), ("deposit_energy", C.c_int),
        ("poisson_tol", C.c_double), ("poisson_max_iter", C.c_int),
        ("mg_pre", C.c_int), ("mg_post", C.c_int), ("mg_omega", C.c_double),
        ("poisson_guess", C.c_int),
        ("poisson_precision", C.c_int), ("poisson_inner_tol", C.c_double),
        ("mg_smoother", C.c_int), ("poisson_direct", C.c_int),
        ("tile_cells", C.c_int), ("deposit_group", C.c_int), ("push_group", C.c_int),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("initial_relres", C.c_double), ("final_relres", C.c_double),
        ("ms", C.c_double), ("guess", C.c_int), ("inner_iterations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Injection(C.Structure):
    _fields_ = [
        ("reinject", C.c_int), ("pairs", C.c_int),
        ("x_inject", C.c_double), ("te_reinject_eV", C.c_double),
        ("te_eV", C.c_double), ("ti_eV", C.c_double),
        ("x11", C.c_double), ("x22", C.c_double), ("xm", C.c_double),
        ("ionimax", C.c_double), ("max_reinject", C.c_int64), ("seed", C.c_uint64),
    ]


_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python bmp-6th-sem_amd/build.py` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sig = {
        "pic_params_default": (None, [C.POINTER(Params)]),
        "pic_create": (i32, [C.POINTER(Params), C.POINTER(vp)]),
        "pic_destroy": (None, [vp]),
        "pic_set_geometry": (i32, [vp, dp, i32]),
        "pic_read_geometry": (i32, [vp, C.c_char_p]),
        "pic_get_geometry": (i32, [vp, ip, dp, C.POINTER(C.c_int)]),
        "pic_add_species": (i32, [vp, dbl, dbl, dbl, dbl, i64, C.POINTER(C.c_int)]),
        "pic_generate_particles": (i32, [C.c_uint64, i32, i64, dbl, dbl, dbl, dbl, ip, dbl, dbl,
                                         dp, dp, dp, dp, dp]),
        "pic_upload_particles": (i32, [vp, i32, i64, dp, dp, dp, dp, dp]),
        "pic_upload_cell32": (i32, [vp, i32, i64, ip, fp, fp, fp, fp, fp]),
        "pic_particle_count": (i64, [vp, i32]),
        "pic_download_particles": (i32, [vp, i32, dp, dp, dp, dp, dp, ip]),
        "pic_download_cell32": (i32, [vp, i32, ip, fp, fp, fp, fp, fp]),
        "pic_bin": (i32, [vp, i32]),
        "pic_get_cell_counts": (i32, [vp, i32, ip]),
        "pic_get_cell_start": (i32, [vp, i32, ip]),
        "pic_zero_rho": (i32, [vp]),
        "pic_deposit": (i32, [vp, i32]),
        "pic_get_deposit_fixed": (i32, [vp, i32, lp, lp]),
        "pic_poisson_solve": (i32, [vp, i32, C.POINTER(SolveStats)]),
        "pic_set_guess": (i32, [vp, dp]),
        "pic_update_fields_from_phi": (i32, [vp]),
        "pic_push": (i32, [vp, i32, i32]),
        "pic_step": (i32, [vp, i32, C.POINTER(SolveStats)]),
        "pic_get_grid": (i32, [vp, i32, dp]),
        "pic_set_grid": (i32, [vp, i32, dp]),
        "pic_get_absorbed": (i32, [vp, i32, lp, lp]),
        "pic_sync": (i32, [vp]),
        "pic_set_timing": (i32, [vp, i32]),
        "pic_stage_times": (i32, [vp, dp, lp]),
        "pic_grid_shape": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "pic_last_error": (C.c_char_p, []),
        "pic_unet_create": (i32, [i32, C.c_uint64, C.POINTER(vp)]),
        "pic_unet_destroy": (None, [vp]),
        "pic_unet_num_params": (i64, [vp]),
        "pic_unet_get_params": (i32, [vp, fp]),
        "pic_unet_set_params": (i32, [vp, fp]),
        "pic_unet_get_grads": (i32, [vp, fp]),
        "pic_unet_forward": (i32, [vp, fp, fp]),
        "pic_unet_train_step": (i32, [vp, fp, fp, C.POINTER(C.c_uint8), C.c_float, fp]),
        "pic_unet_mfma_probe": (i32, [fp, fp, fp]),
        "pic_unet_mfma32_probe": (i32, [fp, fp, fp]),
        "pic_dgemm_check": (i32, [i32, i32, i32, i32, i32, i32, dp, dp, dp, dp]),
        "pic_unet_benchmark": (i32, [vp, i32, fp, fp]),
        "pic_unet_train_step_residual": (i32, [vp, fp, dp, C.POINTER(C.c_uint8), C.c_float, C.c_float, C.c_float,
                                               fp]),
        "pic_learned_unet": (vp, [vp]),
        "pic_ensemble_create": (i32, [C.POINTER(vp), i32, C.POINTER(vp)]),
        "pic_ensemble_destroy": (None, [vp]),
        "pic_ensemble_step": (i32, [vp, i32, i32]),
        "pic_ensemble_poisson_stats": (i32, [vp, ip, dp]),
        "pic_ensemble_tune": (i32, [vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "pic_ensemble_steps_done": (i32, [vp, lp]),
        "pic_learned_enable": (i32, [vp, i32, i32, C.c_float, C.c_uint64]),
        "pic_learned_stats": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), fp]),
        "pic_kinetic_energy": (i32, [vp, i32, dp]),
        "pic_injection_default": (None, [C.POINTER(Injection)]),
        "pic_set_injection": (i32, [vp, C.POINTER(Injection)]),
        "pic_inject": (i32, [vp]),
        "pic_injection_target": (i32, [vp, lp, lp]),
        "pic_injection_stats": (i32, [vp, lp, lp, lp, lp, lp, lp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _chk(rc: int):
    if rc != PIC_OK:
        raise PicError(rc, load().pic_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64)) if a is not None else None


def default_params(**kw) -> Params:
    p = Params()
    load().pic_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def generate_particles(seed: int, stream: int, n: int, width: float, height: float, dx: float, dy: float,
                       qm_abs: float, temp_eV: float, excl=None):
    """Project particle generator (host side of the library)."""
    lib = load()
    out = [np.empty(n, np.float64) for _ in range(5)]
    ex = None
    if excl is not None:
        ex = np.ascontiguousarray(excl, np.int32)
    _chk(lib.pic_generate_particles(seed, stream, n, width, height, dx, dy, _ip(ex), qm_abs, temp_eV,
                                    *[_dp(a) for a in out]))
    return tuple(out)


class Context:
    """One simulation on the GPU (opaque pic_ctx)."""

    def __init__(self, params: Params | None = None, **kw):
        self.lib = load()
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        _chk(self.lib.pic_create(C.byref(self.params), C.byref(h)))
        self.h = h
        gx, gy = C.c_int(), C.c_int()
        _chk(self.lib.pic_grid_shape(self.h, C.byref(gx), C.byref(gy)))
        self.gx, self.gy = gx.value, gy.value
        self.G = self.gx * self.gy
        self.nspecies = 0

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.pic_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dx(self):
        return self.params.width / self.params.NX

    @property
    def dy(self):
        return self.params.height / self.params.NY

    # -- geometry
    def set_geometry(self, rects):
        r = np.ascontiguousarray(rects, np.float64).reshape(-1, 5)
        _chk(self.lib.pic_set_geometry(self.h, _dp(r), r.shape[0]))

    def read_geometry(self, path):
        _chk(self.lib.pic_read_geometry(self.h, str(path).encode()))

    def geometry(self):
        n = C.c_int(4096)
        nodes = np.empty((4096, 4), np.int32)
        pot = np.empty(4096, np.float64)
        _chk(self.lib.pic_get_geometry(self.h, _ip(nodes), _dp(pot), C.byref(n)))
        return nodes[: n.value].copy(), pot[: n.value].copy()

    # -- species / particles
    def add_species(self, charge, mass, sim_count, actual, capacity) -> int:
        sid = C.c_int()
        _chk(self.lib.pic_add_species(self.h, charge, mass, float(sim_count), actual, int(capacity), C.byref(sid)))
        self.nspecies += 1
        return sid.value

    def upload(self, s, x, y, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, vx, vy, vz)]
        _chk(self.lib.pic_upload_particles(self.h, s, arrs[0].size, *[_dp(a) for a in arrs]))

    def upload_cell32(self, s, cell, ox, oy, vx, vy, vz):
        c = np.ascontiguousarray(cell, np.int32)
        f = [np.ascontiguousarray(a, np.float32) for a in (ox, oy, vx, vy, vz)]
        _chk(self.lib.pic_upload_cell32(self.h, s, c.size, _ip(c), *[_fp(a) for a in f]))

    def count(self, s) -> int:
        n = self.lib.pic_particle_count(self.h, s)
        if n < 0:
            _chk(-1)
        return int(n)

    def download(self, s):
        n = self.count(s)
        out = [np.empty(n, np.float64) for _ in range(5)]
        cell = np.empty(n, np.int32)
        _chk(self.lib.pic_download_particles(self.h, s, *[_dp(a) for a in out], _ip(cell)))
        return (*out, cell)

    def download_cell32(self, s):
        n = self.count(s)
        cell = np.empty(n, np.int32)
        f = [np.empty(n, np.float32) for _ in range(5)]
        _chk(self.lib.pic_download_cell32(self.h, s, _ip(cell), *[_fp(a) for a in f]))
        return (cell, *f)

    # -- stages
    def bin(self, s):
        _chk(self.lib.pic_bin(self.h, s))

    def cell_counts(self, s):
        out = np.empty(self.G, np.int32)
        _chk(self.lib.pic_get_cell_counts(self.h, s, _ip(out)))
        return out

    def cell_start(self, s):
        out = np.empty(self.G + 1, np.int32)
        _chk(self.lib.pic_get_cell_start(self.h, s, _ip(out)))
        return out

    def zero_rho(self):
        _chk(self.lib.pic_zero_rho(self.h))

    def deposit(self, s):
        _chk(self.lib.pic_deposit(self.h, s))

    def deposit_fixed(self, s):
        qc = np.empty(self.G, np.int64)
        qe = np.empty(self.G, np.int64)
        _chk(self.lib.pic_get_deposit_fixed(self.h, s, _lp(qc), _lp(qe)))
        return qc, qe

    def poisson_solve(self, iteration=0) -> dict:
        st = SolveStats()
        _chk(self.lib.pic_poisson_solve(self.h, iteration, C.byref(st)))
        return st.as_dict()

    def set_guess(self, phi0):
        a = np.ascontiguousarray(phi0, np.float64)
        _chk(self.lib.pic_set_guess(self.h, _dp(a)))

    def update_fields_from_phi(self):
        _chk(self.lib.pic_update_fields_from_phi(self.h))

    def push(self, s, rebin=True):
        _chk(self.lib.pic_push(self.h, s, 1 if rebin else 0))

    def step(self, iteration) -> dict:
        st = SolveStats()
        _chk(self.lib.pic_step(self.h, iteration, C.byref(st)))
        return st.as_dict()

    def sync(self):
        _chk(self.lib.pic_sync(self.h))

    # -- grids
    def grid(self, which):
        out = np.empty(self.G, np.float64)
        _chk(self.lib.pic_get_grid(self.h, which, _dp(out)))
        return out

    def set_grid(self, which, values):
        a = np.ascontiguousarray(values, np.float64).reshape(-1)
        if a.size != self.G:
            raise ValueError(f"grid needs {self.G} values, got {a.size}")
        _chk(self.lib.pic_set_grid(self.h, which, _dp(a)))

    def absorbed(self, s):
        a, b = C.c_int64(), C.c_int64()
        _chk(self.lib.pic_get_absorbed(self.h, s, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- timing
    def set_timing(self, on=True):
        """0/False off, 1/True per stage, 2 per particle kernel (no markers in the
        stream), 3 the scatter kernel alone."""
        _chk(self.lib.pic_set_timing(self.h, int(on)))

    def learned_enable(self, train_steps, min_trained=0, lr=1e-3, seed=1):
        _chk(self.lib.pic_learned_enable(self.h, int(train_steps), int(min_trained), float(lr), int(seed)))

    def learned_benchmark(self, reps=5):
        """(forward ms, forward + backward ms) of the context's network."""
        h = self.lib.pic_learned_unet(self.h)
        if not h:
            raise PicError(PIC_ERR_STATE, "learned guess not enabled")
        a = np.zeros(1, np.float32)
        b = np.zeros(1, np.float32)
        _chk(self.lib.pic_unet_benchmark(h, int(reps), _fp(a), _fp(b)))
        return float(a[0]), float(b[0])

    def learned_stats(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        l = np.zeros(1, np.float32)
        _chk(self.lib.pic_learned_stats(self.h, C.byref(a), C.byref(b), C.byref(c), _fp(l)))
        return {"trained": a.value, "tried": b.value, "used": c.value, "loss": float(l[0])}

    def kinetic_energy(self, s) -> float:
        """Sum of 0.5 |v|^2 / |q/m| over the live particles of species s."""
        e = np.zeros(1)
        _chk(self.lib.pic_kinetic_energy(self.h, s, _dp(e)))
        return float(e[0])

    # -- injection
    def set_injection(self, enable=True, **kw):
        """Turn injection in step() on with the reference values overridden
        by `kw` (fields of pic_injection), or off."""
        if not enable:
            _chk(self.lib.pic_set_injection(self.h, None))
            return None
        cfg = Injection()
        self.lib.pic_injection_default(C.byref(cfg))
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise KeyError(k)
            setattr(cfg, k, v)
        _chk(self.lib.pic_set_injection(self.h, C.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in cfg._fields_}

    def inject(self):
        _chk(self.lib.pic_inject(self.h))

    def injection_target(self):
        t, a = C.c_int64(), C.c_int64()
        _chk(self.lib.pic_injection_target(self.h, C.byref(t), C.byref(a)))
        return t.value, a.value

    def injection_stats(self):
        v = [C.c_int64() for _ in range(6)]
        _chk(self.lib.pic_injection_stats(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("injections", "reinjected", "pairs", "dropped", "last_electrons", "last_ions"),
                        (x.value for x in v)))

    def stage_times(self):
        ms = np.zeros(len(STAGES), np.float64)
        n = np.zeros(len(STAGES), np.int64)
        _chk(self.lib.pic_stage_times(self.h, _dp(ms), _lp(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(STAGES)}


def mfma_probe(a, b):
    """16x16x16 BF16 MFMA of row-major 16x16 float inputs (exact in BF16)."""
    lib = load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    d = np.empty((16, 16), np.float32)
    _chk(lib.pic_unet_mfma_probe(_fp(a), _fp(b), _fp(d)))
    return d


def mfma32_probe(a, b):
    """16x16x32 BF16 MFMA of row-major a (16 x 32) and b (32 x 16) float inputs (exact in BF16)."""
    lib = load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == (16, 32) and b.shape == (32, 16)
    d = np.empty((16, 16), np.float32)
    _chk(lib.pic_unet_mfma32_probe(_fp(a), _fp(b), _fp(d)))
    return d


def dgemm_check(a, b, ta=False, tb=False, scale=None, c=None):
    """op(a) op(b) on the direct solve's FP64 matrix-core GEMM kernel: op(a) =
    a.T if ta, op(b) = b.T if tb; times `scale` elementwise, or added to `c`."""
    lib = load()
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    M, K = (a.shape[1], a.shape[0]) if ta else a.shape
    K2, N = (b.shape[1], b.shape[0]) if tb else b.shape
    if K != K2:
        raise ValueError(f"inner dimensions differ: {K} vs {K2}")
    ep = 1 if scale is not None else (2 if c is not None else 0)
    out = np.ascontiguousarray(c, np.float64).copy() if c is not None else np.zeros((M, N), np.float64)
    sc = np.ascontiguousarray(scale, np.float64) if scale is not None else None
    _chk(lib.pic_dgemm_check(M, N, K, int(ta), int(tb), ep, _dp(a), _dp(b), _dp(sc) if sc is not None else None,
                             _dp(out)))
    return out


class UNet:
    """The learned-guess network on its own (tests, experiments)."""

    def __init__(self, n: int, seed: int = 1):
        lib = load()
        self.lib = lib
        self.n = n
        h = C.c_void_p()
        _chk(lib.pic_unet_create(n, seed, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.pic_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(self.lib.pic_unet_num_params(self.h))

    def params(self):
        w = np.empty(self.num_params, np.float32)
        _chk(self.lib.pic_unet_get_params(self.h, _fp(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _chk(self.lib.pic_unet_set_params(self.h, _fp(w)))

    def grads(self):
        g = np.empty(self.num_params, np.float32)
        _chk(self.lib.pic_unet_get_grads(self.h, _fp(g)))
        return g

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((self.n, self.n), np.float32)
        _chk(self.lib.pic_unet_forward(self.h, _fp(x), _fp(out)))
        return out

    def train_step_residual(self, x, r, mask, cxd, cyd, lr):
        x = np.ascontiguousarray(x, np.float32)
        r = np.ascontiguousarray(r, np.float64)
        m = np.ascontiguousarray(mask, np.uint8)
        loss = np.zeros(1, np.float32)
        _chk(self.lib.pic_unet_train_step_residual(self.h, _fp(x), _dp(r), m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                   float(cxd), float(cyd), float(lr), _fp(loss)))
        return float(loss[0])

    def train_step(self, x, target, mask, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.uint8)
        loss = np.zeros(1, np.float32)
        _chk(self.lib.pic_unet_train_step(self.h, _fp(x), _fp(t), m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          float(lr), _fp(loss)))
        return float(loss[0])


class Ensemble:
    """Batched steps of several contexts (persistent kernels, one HIP graph)."""

    def __init__(self, contexts):
        self.lib = load()
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        _chk(self.lib.pic_ensemble_create(arr, len(self.contexts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.pic_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, graph=True):
        _chk(self.lib.pic_ensemble_step(self.h, int(n), 1 if graph else 0))

    def poisson_stats(self):
        R = len(self.contexts)
        it = np.zeros(R, np.int32)
        rr = np.zeros(2 * R)
        _chk(self.lib.pic_ensemble_poisson_stats(self.h, _ip(it), _dp(rr)))
        return it, rr.reshape(R, 2)

    def tune(self, steps=2):
        t, b = C.c_int(), C.c_int()
        ms = np.zeros(1)
        _chk(self.lib.pic_ensemble_tune(self.h, int(steps), C.byref(t), C.byref(b), _dp(ms)))
        return {"tile": t.value, "blocks": b.value, "ms_per_step": float(ms[0])}
