"""numpy front end of the C oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check of the graft
entry point and the CPU-baseline leg of bench.py.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build(quiet=True):
    r = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if not quiet:
        print(r.stdout)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    dp, ip, lp, fp = (C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_float))
    u8 = C.POINTER(C.c_uint8)
    sig = {
        "orc_key": (C.c_int32, [dbl, dbl, i32, dbl, dbl]),
        "orc_keys": (None, [i64, dp, dp, i32, dbl, dbl, ip]),
        "orc_bin": (None, [i64, ip, i32, ip, ip, lp]),
        "orc_deposit_ref": (None, [i64, dp, dp, dp, dp, dp, ip, lp, i32, i32, dbl, dbl, dbl, dbl, dbl, dbl,
                                   dp, dp, dp]),
        "orc_deposit_fixed_f64": (None, [i64, dp, dp, dp, dp, dp, ip, i32, i32, dbl, dbl, i32, i32, dbl, lp, lp]),
        "orc_deposit_fixed_cell32": (None, [i64, ip, fp, fp, fp, fp, fp, i32, i32, i32, i32, dbl, lp, lp]),
        "orc_efield": (None, [i32, i32, dbl, dbl, dp, dp, dp]),
        "orc_push_f64": (None, [i64, dp, dp, dp, dp, dp, ip, dp, dp, dp, i32, i32, dbl, dbl, dbl, dbl, dbl, dbl,
                                dbl, dbl, lp]),
        "orc_push_cell32": (None, [i64, ip, fp, fp, fp, fp, fp, ip, dp, dp, dp, i32, i32, dbl, dbl, dbl, dbl, dbl,
                                   dbl, dbl, dbl, lp]),
        "orc_push_f64_p": (None, [i64, dp, dp, dp, dp, dp, ip, dp, dp, dp, i32, i32, dbl, dbl, dbl, dbl, dbl, dbl,
                                  dbl, dbl, i32, lp]),
        "orc_push_cell32_p": (None, [i64, ip, fp, fp, fp, fp, fp, ip, dp, dp, dp, i32, i32, dbl, dbl, dbl, dbl,
                                     dbl, dbl, dbl, dbl, i32, lp]),
        "orc_efield_periodic": (None, [i32, i32, dbl, dbl, dp, dp, dp]),
        "orc_modulo": (dbl, [dbl, dbl]),
        "orc_encode_cell32": (None, [i64, dp, dp, dp, dp, dp, i32, dbl, dbl, ip, fp, fp, fp, fp, fp]),
        "orc_geometry_nodes": (None, [i32, i32, i32, dp, C.POINTER(C.c_int), dp]),
        "orc_dirichlet": (None, [i32, i32, i32, C.POINTER(C.c_int), dp, u8, dp]),
        "orc_overwrite_rho": (None, [i32, u8, dp, dp]),
        "orc_assemble": (i64, [i32, i32, dbl, dbl, u8, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]),
        "orc_relres": (dbl, [i32, i32, dbl, dbl, u8, dp, dp, dp]),
        "orc_poisson_direct": (i32, [i32, i32, dbl, dbl, u8, dp, dp, dp]),
        "orc_have_pardiso": (i32, []),
        "orc_pardiso_create": (vp, [i32, i32, dbl, dbl, u8]),
        "orc_pardiso_solve": (i32, [vp, dp, dp]),
        "orc_pardiso_destroy": (None, [vp]),
        "orc_num_threads": (i32, []),
        "orc_sim_create": (vp, [i32, i32, dbl, dbl, dbl, dbl, dbl, dbl, dbl, i32, dp, i32, i32]),
        "orc_sim_uses_pardiso": (i32, [vp]),
        "orc_sim_add_species": (i32, [vp, dbl, dbl, dbl, dbl, i64, dp, dp, dp, dp, dp]),
        "orc_sim_step": (i32, [vp]),
        "orc_sim_times": (None, [vp, dp]),
        "orc_sim_count": (i64, [vp, i32]),
        "orc_sim_get_grid": (None, [vp, i32, dp]),
        "orc_sim_get_particles": (None, [vp, i32, dp, dp, dp, dp, dp, ip]),
        "orc_sim_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def dp(a):
    return _p(a, C.c_double)


def ip(a):
    return _p(a, C.c_int32)


def lp(a):
    return _p(a, C.c_int64)


def fp(a):
    return _p(a, C.c_float)


def u8(a):
    return _p(a, C.c_uint8)


# ------------------------------------------------------------------ stages
def keys(x, y, gx, dx, dy):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    k = np.empty(x.size, np.int32)
    load().orc_keys(x.size, dp(x), dp(y), gx, dx, dy, ip(k))
    return k


def bin(key, G):
    key = np.ascontiguousarray(key, np.int32)
    counts = np.empty(G, np.int32)
    start = np.empty(G + 1, np.int32)
    order = np.empty(key.size, np.int64)
    load().orc_bin(key.size, ip(key), G, ip(counts), ip(start), lp(order))
    return counts, start, order[: start[-1]]


def deposit_ref(x, y, vx, vy, vz, key, gx, gy, dx, dy, W, qm, charge, eps, rho=None):
    """interpolate restated in FP64: returns (final_mesh, energy_mesh, rho)."""
    G = gx * gy
    arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, vx, vy, vz)]
    counts, start, order = bin(key, G)
    final = np.empty(G)
    energy = np.empty(G)
    rho = np.zeros(G) if rho is None else np.array(rho, np.float64, copy=True)
    load().orc_deposit_ref(arrs[0].size, *[dp(a) for a in arrs], ip(start), lp(order), gx, gy, dx, dy, W, qm,
                           charge, eps, dp(final), dp(energy), dp(rho))
    return final, energy, rho


def deposit_fixed_f64(x, y, vx, vy, vz, key, gx, gy, dx, dy, cs, es, qm):
    G = gx * gy
    arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, vx, vy, vz)]
    key = np.ascontiguousarray(key, np.int32)
    qc = np.empty(G, np.int64)
    qe = np.empty(G, np.int64)
    load().orc_deposit_fixed_f64(arrs[0].size, *[dp(a) for a in arrs], ip(key), gx, gy, dx, dy, cs, es, qm,
                                 lp(qc), lp(qe))
    return qc, qe


def deposit_fixed_cell32(cell, ox, oy, vx, vy, vz, gx, gy, cs, es, qm):
    G = gx * gy
    cell = np.ascontiguousarray(cell, np.int32)
    f = [np.ascontiguousarray(a, np.float32) for a in (ox, oy, vx, vy, vz)]
    qc = np.empty(G, np.int64)
    qe = np.empty(G, np.int64)
    load().orc_deposit_fixed_cell32(cell.size, ip(cell), *[fp(a) for a in f], gx, gy, cs, es, qm, lp(qc), lp(qe))
    return qc, qe


def efield(phi, gx, gy, dx, dy):
    phi = np.ascontiguousarray(phi, np.float64)
    ex = np.zeros(gx * gy)
    ey = np.zeros(gx * gy)
    load().orc_efield(gx, gy, dx, dy, dp(phi), dp(ex), dp(ey))
    return ex, ey


def push_f64(x, y, vx, vy, vz, ex, ey, bz, gx, gy, dx, dy, width, height, qm, dt, bx=0.0, by=0.0,
             periodic=False):
    """Returns the pushed copies (x, y, vx, vy, vz), keys (-1 = absorbed), absorbed [left, total]."""
    a = [np.array(v, np.float64, copy=True) for v in (x, y, vx, vy, vz)]
    f = [np.ascontiguousarray(v, np.float64) for v in (ex, ey, bz)]
    key = np.empty(a[0].size, np.int32)
    ab = np.zeros(2, np.int64)
    load().orc_push_f64_p(a[0].size, *[dp(v) for v in a], ip(key), *[dp(v) for v in f], gx, gy, dx, dy, width,
                          height, qm, dt, bx, by, 1 if periodic else 0, lp(ab))
    return (*a, key, ab)


def push_cell32(cell, ox, oy, vx, vy, vz, ex, ey, bz, gx, gy, dx, dy, width, height, qm, dt, bx=0.0, by=0.0,
                periodic=False):
    c = np.array(cell, np.int32, copy=True)
    a = [np.array(v, np.float32, copy=True) for v in (ox, oy, vx, vy, vz)]
    f = [np.ascontiguousarray(v, np.float64) for v in (ex, ey, bz)]
    key = np.empty(c.size, np.int32)
    ab = np.zeros(2, np.int64)
    load().orc_push_cell32_p(c.size, ip(c), *[fp(v) for v in a], ip(key), *[dp(v) for v in f], gx, gy, dx, dy,
                             width, height, qm, dt, bx, by, 1 if periodic else 0, lp(ab))
    return (c, *a, key, ab)


def modulo(x, y):
    return load().orc_modulo(float(x), float(y))


def efield_periodic(phi, gx, gy, dx, dy):
    phi = np.ascontiguousarray(phi, np.float64)
    ex = np.zeros(gx * gy)
    ey = np.zeros(gx * gy)
    load().orc_efield_periodic(gx, gy, dx, dy, dp(phi), dp(ex), dp(ey))
    return ex, ey


def encode_cell32(x, y, vx, vy, vz, gx, dx, dy):
    a = [np.ascontiguousarray(v, np.float64) for v in (x, y, vx, vy, vz)]
    n = a[0].size
    cell = np.empty(n, np.int32)
    f = [np.empty(n, np.float32) for _ in range(5)]
    load().orc_encode_cell32(n, *[dp(v) for v in a], gx, dx, dy, ip(cell), *[fp(v) for v in f])
    return (cell, *f)


def geometry(gx, gy, rects):
    r = np.ascontiguousarray(rects, np.float64).reshape(-1, 5)
    nodes = np.empty((r.shape[0], 4), np.intc)
    pot = np.empty(r.shape[0])
    load().orc_geometry_nodes(gx, gy, r.shape[0], dp(r), nodes.ctypes.data_as(C.POINTER(C.c_int)), dp(pot))
    mask = np.empty(gx * gy, np.uint8)
    dval = np.empty(gx * gy)
    load().orc_dirichlet(gx, gy, r.shape[0], nodes.ctypes.data_as(C.POINTER(C.c_int)), dp(pot), u8(mask), dp(dval))
    return nodes, pot, mask, dval


def overwrite_rho(rho, mask, dval):
    r = np.array(rho, np.float64, copy=True)
    load().orc_overwrite_rho(r.size, u8(np.ascontiguousarray(mask, np.uint8)), dp(np.ascontiguousarray(dval)),
                             dp(r))
    return r


def assemble(gx, gy, dx, dy, mask):
    G = gx * gy
    ia = np.empty(G + 1, np.intc)
    ja = np.empty(5 * G, np.intc)
    a = np.empty(5 * G)
    nnz = load().orc_assemble(gx, gy, dx, dy, u8(np.ascontiguousarray(mask, np.uint8)),
                              ia.ctypes.data_as(C.POINTER(C.c_int)), ja.ctypes.data_as(C.POINTER(C.c_int)), dp(a))
    return ia, ja[:nnz], a[:nnz]


def relres(gx, gy, dx, dy, mask, rhs, phi, dval):
    return load().orc_relres(gx, gy, dx, dy, u8(np.ascontiguousarray(mask, np.uint8)),
                             dp(np.ascontiguousarray(rhs, np.float64)), dp(np.ascontiguousarray(phi, np.float64)),
                             dp(np.ascontiguousarray(dval, np.float64)))


def poisson_direct(gx, gy, dx, dy, mask, dval, rhs):
    """Solve the reference system for rhs = rho after overwriteRho."""
    phi = np.empty(gx * gy)
    rc = load().orc_poisson_direct(gx, gy, dx, dy, u8(np.ascontiguousarray(mask, np.uint8)),
                                   dp(np.ascontiguousarray(dval, np.float64)),
                                   dp(np.ascontiguousarray(rhs, np.float64)), dp(phi))
    if rc != 0:
        raise RuntimeError(f"orc_poisson_direct failed ({rc})")
    return phi


def poisson_solver(rho, gx, gy, dx, dy, rects):
    """poissonSolver restated: returns (rho_after_overwrite, phi, ex, ey)."""
    _, _, mask, dval = geometry(gx, gy, rects)
    rhs = overwrite_rho(rho, mask, dval)
    phi = poisson_direct(gx, gy, dx, dy, mask, dval, rhs)
    ex, ey = efield(phi, gx, gy, dx, dy)
    return rhs, phi, ex, ey


def assemble_periodic(gx, gy, dx, dy, mask):
    """generateCoeffsPeriodic (src/poissonSolver.cpp:82-180) as a scipy CSR
    matrix of N = gx (gy - 1) rows: Dirichlet rows -1 on the diagonal; other
    rows 2(1/dx^2 + 1/dy^2) on the diagonal, -1/dx^2 at n -+ 1 and -1/dy^2 at
    the rows above and below taken modulo gy - 1.  mask: the first N nodes of
    the geometry mask (nonzero = Dirichlet)."""
    import scipy.sparse as sp

    N = gx * (gy - 1)
    n = np.arange(N)
    d = np.asarray(mask, np.uint8)[:N] != 0
    cx, cy = 1.0 / (dx * dx), 1.0 / (dy * dy)
    u = n[~d]
    rows = [n[d], u, u, u, u, u]
    cols = [n[d], u, u - 1, u + 1, (u - gx) % N, (u + gx) % N]
    vals = [np.full(d.sum(), -1.0), np.full(u.size, 2 * (cx + cy)), np.full(u.size, -cx), np.full(u.size, -cx),
            np.full(u.size, -cy), np.full(u.size, -cy)]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def poisson_periodic(rho, gx, gy, dx, dy, rects):
    """poissonSolver with periodicBoundary = 1, solved by SuperLU (FP64):
    returns (rho after overwriteRho, phi after overwritePhi, ex, ey)."""
    import scipy.sparse.linalg as sla

    _, _, mask, dval = geometry(gx, gy, rects)
    rhs = overwrite_rho(rho, mask, dval)
    N = gx * (gy - 1)
    A = assemble_periodic(gx, gy, dx, dy, mask).tocsr()
    # The Dirichlet rows (-x = rhs) fix x there; eliminating them leaves the
    # SPD system on the unknowns, solved without the 1 : 1/dx^2 row scaling
    # that limits a direct solve of the full matrix.
    d = np.asarray(mask, np.uint8)[:N] != 0
    b = rhs[:N]
    x = np.empty(N)
    x[d] = -b[d]
    AU = A[~d]
    x[~d] = sla.spsolve(AU[:, ~d].tocsc(), b[~d] - AU[:, d] @ x[d])
    phi = np.concatenate([x, x[:gx]])            # overwritePhi: row 0 -> image row
    phi = np.where(mask != 0, dval, phi)         # then the geometry values
    ex, ey = efield_periodic(phi, gx, gy, dx, dy)
    return rhs, phi, ex, ey


class Pardiso:
    def __init__(self, gx, gy, dx, dy, mask):
        self.lib = load()
        if not self.lib.orc_have_pardiso():
            raise RuntimeError("oracle built without MKL PARDISO")
        self.mask = np.ascontiguousarray(mask, np.uint8)
        self.h = self.lib.orc_pardiso_create(gx, gy, dx, dy, u8(self.mask))
        if not self.h:
            raise RuntimeError("PARDISO factorisation failed")
        self.G = gx * gy

    def solve(self, rhs):
        b = np.array(rhs, np.float64, copy=True)
        x = np.empty(self.G)
        rc = self.lib.orc_pardiso_solve(self.h, dp(b), dp(x))
        if rc != 0:
            raise RuntimeError(f"PARDISO phase 33 error {rc}")
        return x

    def close(self):
        if self.h:
            self.lib.orc_pardiso_destroy(self.h)
            self.h = None


class Sim:
    """The CPU baseline simulation (pic_baseline.c)."""

    def __init__(self, NX, NY, width, height, dt, eps, b0, rects, bx=0.0, by=0.0, deposit_energy=1,
                 use_pardiso=True):
        self.lib = load()
        r = np.ascontiguousarray(rects, np.float64).reshape(-1, 5)
        self.h = self.lib.orc_sim_create(NX, NY, width, height, dt, eps, b0, bx, by, r.shape[0], dp(r),
                                         deposit_energy, 1 if use_pardiso else 0)
        self.G = (NX + 1) * (NY + 1)
        self.ns = 0

    def add_species(self, charge, mass, sim_count, actual, x, y, vx, vy, vz):
        a = [np.ascontiguousarray(v, np.float64) for v in (x, y, vx, vy, vz)]
        self.lib.orc_sim_add_species(self.h, charge, mass, float(sim_count), actual, a[0].size, *[dp(v) for v in a])
        self.ns += 1

    @property
    def uses_pardiso(self):
        return bool(self.lib.orc_sim_uses_pardiso(self.h))

    def step(self):
        rc = self.lib.orc_sim_step(self.h)
        if rc != 0:
            raise RuntimeError(f"baseline Poisson solve failed ({rc})")

    def times(self):
        t = np.zeros(5)
        self.lib.orc_sim_times(self.h, dp(t))
        return dict(zip(("deposit", "poisson", "efield", "push", "bin"), t.tolist()))

    def count(self, k):
        return int(self.lib.orc_sim_count(self.h, k))

    def grid(self, which):
        out = np.empty(self.G)
        self.lib.orc_sim_get_grid(self.h, which, dp(out))
        return out

    def particles(self, k):
        n = self.count(k)
        a = [np.empty(n) for _ in range(5)]
        c = np.empty(n, np.int32)
        self.lib.orc_sim_get_particles(self.h, k, *[dp(v) for v in a], ip(c))
        return (*a, c)

    def close(self):
        if self.h:
            self.lib.orc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def threads():
    return int(load().orc_num_threads())
