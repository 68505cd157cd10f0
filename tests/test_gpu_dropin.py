"""The reference's Python driver loop (SYCL-PIC-poissonsolver/src/main.py:75-137)
run through the ``sycl_pic`` mirror module over the drop-in C ABI, against the
oracle's CPU time step on the same particles.

The input files are the reference's own M_PICDAT.DAT layout (the fixture copy
in tests/golden) with the grid, time step, field and species rows replaced by
a 64x64 case, and a geometry.conf with an interior electrode at +100 V.
"""
import numpy as np
import pytest

from conftest import GOLDEN

E, ME, MI = 1.602176634e-19, 9.1093837015e-31, 40 * 1.66053906660e-27
NX, N, DX, DT, B0_T = 64, 20000, 1e-4, 1.77e-11, 0.01
RECTS = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0],
         [0.4375, 0.4375, 0.5625, 0.5625, 100.0]]


def _picdat() -> str:
    lines = (GOLDEN / "M_PICDAT_reference.DAT").read_text().splitlines()
    W = NX * DX

    def put(header_start, value, offset=1):
        i = next(k for k, s in enumerate(lines) if s.startswith(header_start))
        lines[i + offset] = value

    put("TIMESTEPS", "3")
    put("XMAX(m)", f"{W!r}\t\t{W!r}\t\t{NX}\t\t{NX}")
    put("dt(s)", f"{DT!r}")
    put("B0(mT)", f"{B0_T * 1000!r}\t\t0.02\t\t0.7")
    put("Te(eV)", "2\t\t0.03\t\t300")
    put("speciesId", f"1\t{N}\t1e16\t{-E!r}\t{ME!r}")
    put("speciesId", f"2\t{N}\t1e16\t{E!r}\t{MI!r}", offset=2)
    return "\n".join(lines) + "\n"


def _geometry() -> str:
    rows = "\n".join(" ".join(repr(float(v)) for v in r) for r in RECTS)
    return f"header\nheader\n{len(RECTS)}\nheader\nx1 y1 x2 y2 V\n{rows}\n"


@pytest.mark.gpu
def test_python_driver_loop_matches_oracle(hip_mod, oracle_mod, tmp_path, monkeypatch):
    hip, O = hip_mod, oracle_mod
    (tmp_path / "M_PICDAT.DAT").write_text(_picdat())
    (tmp_path / "geometry.conf").write_text(_geometry())
    for d in ("electronDensity", "ionDensity", "electronenergy", "ionenergy", "Ex", "Ey", "potential"):
        (tmp_path / "out" / d).mkdir(parents=True)
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("BMP_PIC_SEED", "1")
    import sycl_pic

    W = NX * DX
    sim = O.Sim(NX, NX, W, W, DT, 8.854e-12, B0_T, RECTS, deposit_energy=1, use_pardiso=True)
    for s, (q, m, T) in enumerate([(-E, ME, 2.0), (E, MI, 0.03)]):
        parts = hip.generate_particles(1, s, N, W, W, DX, DX, abs(q / m), T)
        sim.add_species(q, m, N, 1e16, *parts)

    sycl_pic.init("gpu", 256, 8, "electronParticle.inp", "ionParticle.inp")
    try:
        assert sycl_pic.get_grid_x() == NX + 1 and sycl_pic.get_grid_y() == NX + 1
        assert sycl_pic.get_grid_size() == (NX + 1) ** 2
        assert sycl_pic.get_timesteps() == 3 and sycl_pic.get_num_species() == 2
        assert sycl_pic.get_particle_counts() == (N, N)
        rho, phi = sycl_pic.get_rho(), sycl_pic.get_phi()
        assert rho.shape == ((NX + 1) ** 2,) and not rho.flags.owndata
        pi = sycl_pic.get_print_interval()
        for it in range(sycl_pic.get_timesteps()):
            sycl_pic.zero_rho()
            for sp in range(2):
                assert sycl_pic.charge_deposition(sp) >= 0
            sycl_pic.poisson_solve(it)
            for sp in range(2):
                sycl_pic.move_particles(sp, it)
                if it % pi == 0:
                    t, a, b = sycl_pic.compact_particles(sp)
                    assert t >= 0
                sycl_pic.cleanup_map(sp)
            if it % pi == 0:
                sycl_pic.print_diagnostics("./out", it)
            sim.step()
            for name, a, which in (("rho", rho, hip.GRID_RHO), ("phi", phi, hip.GRID_PHI)):
                b = sim.grid(which)
                err = np.abs(a - b).max() / np.abs(b).max()
                assert err < 1e-9, (it, name, err)
            ne, ni = sycl_pic.get_particle_counts()
            assert abs(ne - sim.count(0)) <= 2 and abs(ni - sim.count(1)) <= 2
        # diagnostics text grids: the potential file holds phi, "%lf" rows
        pot = np.loadtxt(tmp_path / "out" / "potential" / "0")
        assert pot.shape == (NX + 1, NX + 1)
        ex = np.loadtxt(tmp_path / "out" / "Ex" / "2").ravel()
        ref_ex = sim.grid(hip.GRID_EX)
        # "%lf" keeps 6 decimals; phi agrees to 1e-9 relative, E to about phi's error / dx
        assert np.abs(ex - ref_ex).max() <= 1e-6 + 1e-7 * np.abs(ref_ex).max()

        # ML hook: a caller writes phi through the view, then asks for the fields
        gx = NX + 1
        jj = np.tile(np.arange(gx), gx)
        phi[:] = 3.0 * jj * DX
        sycl_pic.update_fields_from_phi()
        p2 = phi.reshape(gx, gx)
        assert p2[0, 5] == 0.0 and p2[gx // 2, gx // 2] == 100.0   # boundary values rewritten
        sycl_pic.print_diagnostics("./out", 99)
        ex2 = np.loadtxt(tmp_path / "out" / "Ex" / "99")
        # central difference of a ramp of slope 3 V/m away from the masks
        assert np.allclose(ex2[10:20, 10:20], -3.0, atol=1e-6)
    finally:
        sycl_pic.finalize()
        sim.close()


@pytest.mark.gpu
def test_driver_loop_with_generate_new_particles(hip_mod, tmp_path, monkeypatch):
    """main.cpp's loop order (movers, then generate_new_particles): each call
    adds this step's ionization pairs to both species; the re-injection plane
    x = 0.024 m lies outside this 6.4 mm domain, so re-injected electrons are
    dropped as the reference's out-of-grid insert drops them (utils.cpp:456-461)."""
    import inject_oracle as IO

    (tmp_path / "M_PICDAT.DAT").write_text(_picdat())
    (tmp_path / "geometry.conf").write_text(_geometry())
    monkeypatch.chdir(tmp_path)
    import sycl_pic

    target, _ = IO.pair_target(NX * DX, DT)
    sycl_pic.init("gpu", 256, 8, "electronParticle.inp", "ionParticle.inp")
    try:
        for it in range(3):
            sycl_pic.zero_rho()
            for sp in range(2):
                sycl_pic.charge_deposition(sp)
            sycl_pic.poisson_solve(it)
            for sp in range(2):
                sycl_pic.move_particles(sp, it)
            ne, ni = sycl_pic.get_particle_counts()
            assert sycl_pic.generate_new_particles() >= 0
            assert sycl_pic.get_particle_counts() == (ne + target, ni + target)
    finally:
        sycl_pic.finalize()


def test_module_surface_matches_reference_table(hip_mod):
    """Every method of the reference module table (sycl_pic_module.cpp:386-410)."""
    import sycl_pic

    names = ["init", "get_grid_size", "get_grid_x", "get_grid_y", "get_timesteps", "get_print_interval",
             "get_num_species", "get_particle_counts", "get_rho", "get_phi", "zero_rho", "charge_deposition",
             "poisson_solve", "update_fields_from_phi", "move_particles", "compact_particles", "cleanup_map",
             "generate_new_particles", "print_diagnostics", "save_meshes", "dump_rho_phi_cpp", "finalize"]
    for n in names:
        assert callable(getattr(sycl_pic, n)), n
    with pytest.raises(ValueError):
        sycl_pic.init("cpu", 256, 8, "a", "b")
    with pytest.raises(ValueError):
        sycl_pic.init("fpga", 256, 8, "a", "b")
