"""Python surface of the reference's ``sycl_pic`` extension module, over the
HIP library's drop-in C ABI.

Same function names, arguments and return values as the module table in
SYCL-PIC-poissonsolver/src/sycl_pic_module.cpp:386-410, so a driver written
for it (SYCL-PIC-poissonsolver/src/main.py:75-137) runs unchanged with
``import sycl_pic`` resolving here:

* ``get_rho()`` / ``get_phi()`` return NumPy views (no copy, ``OWNDATA``
  false) of the library's pinned host mirrors of the device grids, valid until
  ``finalize()``.  Writes made through them (the in-place ML solve of
  SYCL-PIC-poissonsolver/src/poisson_ml_bridge.py:55-106) are copied to the
  device by the next call that reads that grid (``poisson_solve`` for rho,
  ``update_fields_from_phi`` for phi); calls that change a grid refresh its
  mirror before they return.
* Timed calls return wall seconds of a synchronous call.
* ``init`` accepts the reference's two device values (sycl_pic_module.cpp:63-71);
  both run the simulation on the GPU, the only device this library drives
  (``"cpu"`` prints a note saying so).  Any other value raises
  ``ValueError``; missing or malformed input files raise ``RuntimeError``.

The library must have been built (``python bmp-6th-sem_amd/build.py``);
importing this module on a machine without it raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import pic_hip

_lib = pic_hip.load()
_dbl, _int = C.c_double, C.c_int
_sig = {
    "sycl_pic_init": (_int, [C.c_char_p, _int, _int, C.c_char_p, C.c_char_p]),
    "sycl_pic_get_grid_size": (_int, []),
    "sycl_pic_get_grid_x": (_int, []),
    "sycl_pic_get_grid_y": (_int, []),
    "sycl_pic_get_timesteps": (_int, []),
    "sycl_pic_get_print_interval": (_int, []),
    "sycl_pic_get_num_species": (_int, []),
    "sycl_pic_get_particle_count": (_int, [_int]),
    "sycl_pic_get_rho_ptr": (C.POINTER(_dbl), []),
    "sycl_pic_get_phi_ptr": (C.POINTER(_dbl), []),
    "sycl_pic_zero_rho": (None, []),
    "sycl_pic_charge_deposition": (_dbl, [_int]),
    "sycl_pic_poisson_solve": (_dbl, [_int]),
    "sycl_pic_update_fields_from_phi": (None, []),
    "sycl_pic_move_particles": (_dbl, [_int, _int]),
    "sycl_pic_compact_particles": (_dbl, [_int, C.POINTER(_dbl), C.POINTER(_dbl)]),
    "sycl_pic_cleanup_map": (_dbl, [_int]),
    "sycl_pic_generate_new_particles": (_dbl, []),
    "sycl_pic_print_diagnostics": (None, [C.c_char_p, _int]),
    "sycl_pic_save_meshes": (None, []),
    "sycl_pic_finalize": (None, []),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

_views: dict[str, np.ndarray] = {}


def _view(ptr) -> np.ndarray:
    n = _lib.sycl_pic_get_grid_size()
    if not ptr or n <= 0:
        raise RuntimeError("sycl_pic is not initialised")
    return np.ctypeslib.as_array(ptr, shape=(n,))


def init(device: str, wg_size: int, prob_sz: int, input1: str, input2: str) -> None:
    """Read ./M_PICDAT.DAT and ./geometry.conf, create the particles and the
    Poisson hierarchy (sycl_pic_module.cpp:54-93)."""
    if device not in ("cpu", "gpu"):
        raise ValueError("device must be 'cpu' or 'gpu'")
    rc = _lib.sycl_pic_init(device.encode(), int(wg_size), int(prob_sz), input1.encode(), input2.encode())
    if rc != 0:
        raise RuntimeError(f"sycl_pic.init failed ({rc}): {_lib.pic_last_error().decode(errors='replace')}")


def get_grid_size() -> int:
    return _lib.sycl_pic_get_grid_size()


def get_grid_x() -> int:
    return _lib.sycl_pic_get_grid_x()


def get_grid_y() -> int:
    return _lib.sycl_pic_get_grid_y()


def get_timesteps() -> int:
    return _lib.sycl_pic_get_timesteps()


def get_print_interval() -> int:
    return _lib.sycl_pic_get_print_interval()


def get_num_species() -> int:
    return _lib.sycl_pic_get_num_species()


def get_particle_counts() -> tuple[int, int]:
    """(electrons, ions) live counts (sycl_pic_module.cpp:145-149 returns the
    high-water marks; after a sort there are no holes, so they coincide)."""
    return _lib.sycl_pic_get_particle_count(0), _lib.sycl_pic_get_particle_count(1)


def get_rho() -> np.ndarray:
    """rho as a flat float64 view of GRID_X*GRID_Y nodes, row-major [gy*GRID_X+gx]."""
    if "rho" not in _views:
        _views["rho"] = _view(_lib.sycl_pic_get_rho_ptr())
    return _views["rho"]


def get_phi() -> np.ndarray:
    if "phi" not in _views:
        _views["phi"] = _view(_lib.sycl_pic_get_phi_ptr())
    return _views["phi"]


def zero_rho() -> None:
    _lib.sycl_pic_zero_rho()


def charge_deposition(species: int) -> float:
    return _lib.sycl_pic_charge_deposition(int(species))


def poisson_solve(iteration: int) -> float:
    return _lib.sycl_pic_poisson_solve(int(iteration))


def update_fields_from_phi() -> None:
    _lib.sycl_pic_update_fields_from_phi()


def move_particles(species: int, iteration: int) -> float:
    return _lib.sycl_pic_move_particles(int(species), int(iteration))


def compact_particles(species: int) -> tuple[float, float, float]:
    """(elapsed, alloc, scan) seconds (sycl_pic_module.cpp:247-263)."""
    a, s = _dbl(0.0), _dbl(0.0)
    t = _lib.sycl_pic_compact_particles(int(species), C.byref(a), C.byref(s))
    return t, a.value, s.value


def cleanup_map(species: int) -> float:
    return _lib.sycl_pic_cleanup_map(int(species))


def generate_new_particles() -> float:
    return _lib.sycl_pic_generate_new_particles()


def print_diagnostics(directory: str, iteration: int) -> None:
    _lib.sycl_pic_print_diagnostics(directory.encode(), int(iteration))


def save_meshes() -> None:
    _lib.sycl_pic_save_meshes()


def dump_rho_phi_cpp(prefix: str) -> None:
    """Write rho and phi as raw float64 to <prefix>_rho_cpp.bin and
    <prefix>_phi_cpp.bin (sycl_pic_module.cpp:350-372)."""
    n = get_grid_size()
    for name, arr in (("rho", get_rho()), ("phi", get_phi())):
        path = f"{prefix}_{name}_cpp.bin"
        np.ascontiguousarray(arr, np.float64).tofile(path)
        print(f"[CPP-DUMP] Wrote {name} ({n} doubles) to {path}")


def finalize() -> None:
    _views.clear()
    _lib.sycl_pic_finalize()
