"""The C-ABI library loads without a GPU and exports every symbol the public
header declares; host-only entry points behave as documented."""
import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "bmp_pic.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pic_[a-z0-9_]+|sycl_pic_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_drop_in_surface():
    names = declared_functions()
    # the reference's plain C wrapper (sycl_pic_ctypes.cpp:46-272)
    for n in ["sycl_pic_init", "sycl_pic_get_grid_size", "sycl_pic_get_timesteps", "sycl_pic_get_print_interval",
              "sycl_pic_get_num_species", "sycl_pic_get_particle_count", "sycl_pic_get_rho_ptr",
              "sycl_pic_get_phi_ptr", "sycl_pic_zero_rho", "sycl_pic_charge_deposition", "sycl_pic_poisson_solve",
              "sycl_pic_move_particles", "sycl_pic_compact_particles", "sycl_pic_cleanup_map",
              "sycl_pic_generate_new_particles", "sycl_pic_print_diagnostics", "sycl_pic_save_meshes",
              "sycl_pic_finalize", "sycl_pic_update_fields_from_phi"]:
        assert n in names, n


def test_library_exports_every_declared_symbol(hip_mod):
    lib = hip_mod.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_drop_in_device_values(hip_mod, tmp_path, monkeypatch):
    """sycl_pic_init takes the reference's two device values ("cpu" and "gpu",
    sycl_pic_ctypes.cpp:53-60) and returns -1 for any other; in a directory
    without M_PICDAT.DAT both accepted values go on to report the missing file
    (PIC_ERR_FILE) before anything touches a device."""
    lib = hip_mod.load()
    lib.sycl_pic_init.restype = ctypes.c_int
    lib.sycl_pic_init.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p]
    monkeypatch.chdir(tmp_path)
    assert lib.sycl_pic_init(b"fpga", 256, 8, b"a", b"b") == -1
    assert lib.sycl_pic_init(b"", 256, 8, b"a", b"b") == -1
    assert lib.sycl_pic_init(b"cpu", 256, 8, b"a", b"b") == -2
    assert lib.sycl_pic_init(b"gpu", 256, 8, b"a", b"b") == -2


def test_generator_is_deterministic_and_maxwellian(hip_mod):
    """Counter-based generator: same seed -> same particles independent of n;
    velocity components have sigma = sqrt(|q/m| T)."""
    qm, T = 1.758820e11, 2.0
    a = hip_mod.generate_particles(3, 0, 200000, 0.1, 0.05, 1e-4, 1e-4, qm, T)
    b = hip_mod.generate_particles(3, 0, 1000, 0.1, 0.05, 1e-4, 1e-4, qm, T)
    for u, v in zip(a, b):
        assert np.array_equal(u[:1000], v)
    x, y, vx, vy, vz = a
    assert x.min() >= 0 and x.max() < 0.1 and y.min() >= 0 and y.max() < 0.05
    sig = np.sqrt(qm * T)
    for v in (vx, vy, vz):
        assert abs(v.std() / sig - 1) < 0.01 and abs(v.mean()) < 0.01 * sig
    c = hip_mod.generate_particles(4, 0, 1000, 0.1, 0.05, 1e-4, 1e-4, qm, T)
    assert not np.array_equal(c[0], b[0])


def test_generator_excludes_electrode(hip_mod):
    excl = [448, 448, 576, 576]
    x, y, *_ = hip_mod.generate_particles(3, 1, 100000, 0.1024, 0.1024, 1e-4, 1e-4, 1e6, 0.03, excl=excl)
    inside = (x >= 448e-4) & (x <= 576e-4) & (y >= 448e-4) & (y <= 576e-4)
    assert not inside.any()


def test_context_needs_gpu(hip_mod):
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(hip_mod.PicError) as e:
        hip_mod.Context(hip_mod.default_params())
    assert e.value.code == -10
