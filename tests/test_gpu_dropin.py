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
        sycl_pic.init("fpga", 256, 8, "a", "b")
    # the C entry point returns -1 for an unknown device (sycl_pic_ctypes.cpp:57-59)
    assert sycl_pic._lib.sycl_pic_init(b"fpga", 256, 8, b"a", b"b") == -1


@pytest.mark.gpu
def test_diagnostics_files_match_oracle(hip_mod, oracle_mod, tmp_path, monkeypatch):
    """Every file the reference's diagnostics write, read back with loadtxt and
    compared with the oracle's time step on the same particles, under the
    reference's formulas (SYCL-PIC-poissonsolver/src/utils.cpp:829-1025,
    src/main.cpp:260-274):

    * electronDensity / ionDensity: final_mesh * scaleFactor (scale 3 here);
    * electronenergy / ionenergy: 0.66667 energy / (final + 1), except the last
      electron column, which divides by energy + 1 (utils.cpp:914);
    * Ex, Ey, potential: the fields;
    * electronNumber / ionNumber (written by the mover, mover.cpp:421-425):
      per-key particle counts after the push, exact integers;
    * Mesh_<s>.out: final_mesh through an ostream (6 significant digits).

    "%lf" keeps 6 decimals, so values compare to 1e-9 relative plus 1e-6."""
    hip, O = hip_mod, oracle_mod
    text = _picdat().replace("Scale factor\n1\n", "Scale factor\n3\n")
    assert "Scale factor\n3\n" in text
    (tmp_path / "M_PICDAT.DAT").write_text(text)
    (tmp_path / "geometry.conf").write_text(_geometry())
    dirs = ("electronDensity", "ionDensity", "electronenergy", "ionenergy", "Ex", "Ey", "potential",
            "electronNumber", "ionNumber")
    for d in dirs:
        (tmp_path / "out" / d).mkdir(parents=True)
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("BMP_PIC_SEED", "1")
    import sycl_pic

    W = NX * DX
    gx = NX + 1
    sim = O.Sim(NX, NX, W, W, DT, 8.854e-12, B0_T, RECTS, deposit_energy=1, use_pardiso=True)
    for s, (q, m, T) in enumerate([(-E, ME, 2.0), (E, MI, 0.03)]):
        parts = hip.generate_particles(1, s, N, W, W, DX, DX, abs(q / m), T)
        sim.add_species(q, m, N, 1e16, *parts)

    def close(path, want, rel=1e-9):
        got = np.loadtxt(path)
        assert got.shape == (gx, gx), path
        want = want.reshape(gx, gx)
        err = np.abs(got - want).max()
        assert err <= 1e-6 + rel * np.abs(want).max(), (path, err)

    # the reference's other device value runs the same GPU path
    sycl_pic.init("cpu", 256, 8, "electronParticle.inp", "ionParticle.inp")
    try:
        for it in range(2):
            sycl_pic.zero_rho()
            for sp in range(2):
                sycl_pic.charge_deposition(sp)
            sycl_pic.poisson_solve(it)
            for sp in range(2):
                sycl_pic.move_particles(sp, it)
            sycl_pic.print_diagnostics("./out", it)
            sim.step()
            out = tmp_path / "out"
            fm = [sim.grid(hip.GRID_FINAL0 + s) for s in range(2)]
            em = [sim.grid(hip.GRID_ENERGY0 + s) for s in range(2)]
            close(out / "electronDensity" / str(it), 3.0 * fm[0])
            close(out / "ionDensity" / str(it), 3.0 * fm[1])
            e_want = 0.66667 * em[0] / (fm[0] + 1.0)
            quirk = (0.66667 * em[0] / (em[0] + 1.0)).reshape(gx, gx)[:, -1]
            e_want = e_want.reshape(gx, gx).copy()
            e_want[:, -1] = quirk
            close(out / "electronenergy" / str(it), e_want)
            close(out / "ionenergy" / str(it), 0.66667 * em[1] / (fm[1] + 1.0))
            # the quirk is visible: the last electron column is not energy / density
            plain = (0.66667 * em[0] / (fm[0] + 1.0)).reshape(gx, gx)[:, -1]
            if np.abs(plain - quirk).max() > 1e-3:
                got_last = np.loadtxt(out / "electronenergy" / str(it))[:, -1]
                assert np.abs(got_last - quirk).max() < np.abs(got_last - plain).max()
            close(out / "Ex" / str(it), sim.grid(hip.GRID_EX))
            close(out / "Ey" / str(it), sim.grid(hip.GRID_EY))
            close(out / "potential" / str(it), sim.grid(hip.GRID_PHI))
            for s, name in enumerate(("electronNumber", "ionNumber")):
                got = np.loadtxt(out / name / str(it), dtype=np.int64)
                want = np.bincount(sim.particles(s)[5], minlength=gx * gx).reshape(gx, gx)
                assert got.shape == (gx, gx)
                assert np.array_equal(got, want), (name, it, int(np.abs(got - want).sum()))
                assert got.sum() == sycl_pic.get_particle_counts()[s]
        sycl_pic.save_meshes()
        for s in range(2):
            got = np.loadtxt(tmp_path / f"Mesh_{s}.out")
            want = sim.grid(hip.GRID_FINAL0 + s).reshape(gx, gx)
            assert got.shape == (gx, gx)
            assert np.allclose(got, want, rtol=6e-6, atol=0), s
    finally:
        sycl_pic.finalize()
        sim.close()
