"""Periodic y (periodicBoundary = 1) on the GPU against the oracle: the
Poisson solve against a direct solve of the reference's periodic system
(generateCoeffsPeriodic, poissonSolver.cpp:82-180) to 1e-10 relative, the
periodic E field (:411-430) bit for bit on the same phi, the push with the
mover's y wrap (mover.cpp:13-22,317-320) bit for bit (F64) / against the
CELL32 restatement, and three full steps against the oracle's stages."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, ME, MI = 1.602176634e-19, 9.1093837015e-31, 40 * 1.66053906660e-27
PER_GEOM = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0.4375, 0.4375, 0.5625, 0.5625, 100.0]]


def _ctx(hip, NX, layout, NY=None, **kw):
    NY = NY or NX
    dx = 1e-4
    p = hip.default_params(NX=NX, NY=NY, width=NX * dx, height=NY * dx, dt=1.77e-11, b0=0.01, layout=layout,
                           periodic=1, **kw)
    c = hip.Context(p)
    c.set_geometry(PER_GEOM)
    return c


def _canon(*cols):
    a = np.stack([np.asarray(v, np.float64) for v in cols], 1)
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("NX,NY,precision", [(64, 64, 0), (128, 64, 0), (128, 128, 1)])
def test_periodic_poisson_and_efield(hip_mod, oracle_mod, NX, NY, precision):
    hip, O = hip_mod, oracle_mod
    c = _ctx(hip, NX, 0, NY=NY, poisson_tol=1e-12, poisson_precision=precision)
    rng = np.random.default_rng(NX + NY)
    rho = rng.normal(0, 1e6, c.G)
    c.set_grid(hip.GRID_RHO, rho)
    st = c.poisson_solve(0)
    assert st["final_relres"] <= 1e-12
    rhs, phi_ref, _, _ = O.poisson_periodic(rho, c.gx, c.gy, c.dx, c.dy, PER_GEOM)
    phi = c.grid(hip.GRID_PHI)
    assert np.abs(phi - phi_ref).max() / np.abs(phi_ref).max() < 1e-10
    p2 = phi.reshape(c.gy, c.gx)
    assert np.array_equal(p2[-1], p2[0])
    assert np.array_equal(c.grid(hip.GRID_RHO), rhs)        # overwriteRho in place
    ex, ey = O.efield_periodic(phi, c.gx, c.gy, c.dx, c.dy)
    assert np.array_equal(c.grid(hip.GRID_EX), ex)
    assert np.array_equal(c.grid(hip.GRID_EY), ey)
    c.close()


@pytest.mark.parametrize("layout", [0, 1])
def test_periodic_push_parity(hip_mod, oracle_mod, layout):
    hip, O = hip_mod, oracle_mod
    NX = 64
    c = _ctx(hip, NX, layout)
    rng = np.random.default_rng(5)
    ex, ey = rng.normal(0, 2e3, (2, c.G))
    bz = 0.0433 + rng.normal(0, 1e-3, c.G)
    for which, g in ((hip.GRID_EX, ex), (hip.GRID_EY, ey), (hip.GRID_BZ, bz)):
        c.set_grid(which, g)
    W, H, dt = c.params.width, c.params.height, c.params.dt
    g = dict(gx=c.gx, gy=c.gy, dx=c.dx, dy=c.dy, width=W, height=H)
    for sp, (q, m, T) in enumerate([(-E, ME, 2.0), (E, MI, 0.03)]):
        n = 40000
        s = c.add_species(q, m, n, 1e16, n)
        x, y, vx, vy, vz = hip.generate_particles(9, sp, n, W, H, c.dx, c.dy, abs(q / m), T)
        # a third of them next to the top or bottom row, moving out through it
        k = n // 3
        y[:k] = np.where(np.arange(k) % 2 == 0, H - 0.3 * c.dy * rng.random(k), 0.3 * c.dy * rng.random(k))
        vy[:k] = np.where(np.arange(k) % 2 == 0, 1.0, -1.0) * c.dy / dt * (0.5 + rng.random(k))
        c.upload(s, x, y, vx, vy, vz)
        if layout == 0:
            X, Y, VX, VY, VZ, key, ab = O.push_f64(x, y, vx, vy, vz, ex, ey, bz, qm=q / m, dt=dt, periodic=True, **g)
        else:
            cell, ox, oy, fvx, fvy, fvz = O.encode_cell32(x, y, vx, vy, vz, c.gx, c.dx, c.dy)
            rc, rox, roy, rvx, rvy, rvz, key, ab = O.push_cell32(cell, ox, oy, fvx, fvy, fvz, ex, ey, bz, qm=q / m,
                                                                 dt=dt, periodic=True, **g)
        c.push(s, rebin=True)
        live = key >= 0
        assert c.count(s) == live.sum()
        assert c.absorbed(s) == (ab[0], ab[1])
        # nothing leaves through y: the absorbed ones crossed an x wall
        if layout == 0:
            gone = key[:k] < 0
            assert ((X[:k][gone] < 0) | (X[:k][gone] >= W)).all()
            assert gone.sum() < k // 50
        if layout == 0:
            gx_, gy_, gvx, gvy, gvz, gcell = c.download(s)
            assert np.array_equal(_canon(gcell, gx_, gy_, gvx, gvy, gvz),
                                  _canon(key[live], X[live], Y[live], VX[live], VY[live], VZ[live]))
        else:
            cell2, ox2, oy2, v1, v2, v3 = c.download_cell32(s)
            assert np.array_equal(_canon(cell2, ox2, oy2, v1, v2, v3),
                                  _canon(key[live], rox[live], roy[live], rvx[live], rvy[live], rvz[live]))
    c.close()


def test_periodic_steps_against_oracle_stages(hip_mod, oracle_mod):
    """pic_step with periodic y: each step's rho (CIC deposit, 1e-12), phi
    (periodic direct solve, 1e-9) and pushed particles (bit for bit with the
    device's E) against the oracle."""
    hip, O = hip_mod, oracle_mod
    NX = 64
    c = _ctx(hip, NX, 0, poisson_tol=1e-13)
    W, H, dt = c.params.width, c.params.height, c.params.dt
    n = 20000
    qm = []
    for sp, (q, m, T) in enumerate([(-E, ME, 2.0), (E, MI, 0.03)]):
        s = c.add_species(q, m, n, 1e16, n)
        parts = hip.generate_particles(3, sp, n, W, H, c.dx, c.dy, abs(q / m), T)
        c.upload(s, *parts)
        qm.append((q, m))
    g = dict(gx=c.gx, gy=c.gy, dx=c.dx, dy=c.dy, width=W, height=H)
    for it in range(3):
        state = [c.download(s) for s in range(2)]
        c.step(it)
        rho = np.zeros(c.G)
        for s, (q, m) in enumerate(qm):
            x, y, vx, vy, vz, cell = state[s]
            w = 1e16 * (W * H) / (n * c.dx * c.dy * c.dx * c.dy)
            _, _, rho = O.deposit_ref(x, y, vx, vy, vz, cell, c.gx, c.gy, c.dx, c.dy, w, q / m, q, c.params.epsilon,
                                      rho)
        rhs, phi_ref, ex_ref, ey_ref = O.poisson_periodic(rho, c.gx, c.gy, c.dx, c.dy, PER_GEOM)
        assert np.abs(c.grid(hip.GRID_RHO) - rhs).max() <= 1e-12 * np.abs(rhs).max()
        phi = c.grid(hip.GRID_PHI)
        assert np.abs(phi - phi_ref).max() <= 1e-9 * np.abs(phi_ref).max()
        ex, ey, bz = c.grid(hip.GRID_EX), c.grid(hip.GRID_EY), c.grid(hip.GRID_BZ)
        for s, (q, m) in enumerate(qm):
            x, y, vx, vy, vz, _ = state[s]
            X, Y, VX, VY, VZ, key, _ = O.push_f64(x, y, vx, vy, vz, ex, ey, bz, qm=q / m, dt=dt, periodic=True,
                                                  **g)
            live = key >= 0
            got = c.download(s)
            assert np.array_equal(_canon(got[5], *got[:5]),
                                  _canon(key[live], X[live], Y[live], VX[live], VY[live], VZ[live]))
    c.close()


@pytest.mark.parametrize("solver", ["pcg", "direct"])
@pytest.mark.parametrize("NX,NY,periodic", [(500, 256, 1), (500, 256, 0), (600, 600, 0), (97, 61, 0)])
def test_grids_with_odd_cell_counts(hip_mod, oracle_mod, NX, NY, periodic, solver):
    """The multigrid hierarchy coarsens levels with an odd cell count by one
    overhanging Dirichlet coarse node, so the PCG alone solves grids whose
    sizes have few factors of two: the reference's published grid (501 x 257
    nodes, periodic y, Manya_Vansh_BMP.md:155-167) and non-periodic grids of
    500 x 256, 600 x 600 and 97 x 61 cells with an electrode.  Against the
    oracle (SuperLU on the periodic matrix; MKL PARDISO or the band Cholesky
    on the Dirichlet one) to 1e-10 at tolerance 1e-12.  The same grids by the
    direct solve, whose transforms are GEMMs here (no size is 2^k - 1): no PCG
    iteration left to do."""
    hip, O = hip_mod, oracle_mod
    dx = 1e-4
    elec = [0.4375, 0.4375, 0.5625, 0.5625, 100.0]
    walls = [[0, 0, 0, 1, 200.0], [1, 0, 1, 1, 0.0]]
    geom = walls + [elec] if periodic else walls + [[0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0], elec]
    p = hip.default_params(NX=NX, NY=NY, width=NX * dx, height=NY * dx, dt=1.77e-11, layout=0, periodic=periodic,
                           poisson_tol=1e-12, poisson_direct=1 if solver == "direct" else 0)
    c = hip.Context(p)
    c.set_geometry(geom)
    rng = np.random.default_rng(NX * NY)
    rho = rng.normal(0, 1e6, c.G)
    c.set_grid(hip.GRID_RHO, rho)
    st = c.poisson_solve(0)
    assert st["final_relres"] <= 1e-12, st
    if periodic:
        _, phi_ref, _, _ = O.poisson_periodic(rho, c.gx, c.gy, c.dx, c.dy, geom)
    else:
        _, _, mask, dval = O.geometry(c.gx, c.gy, geom)
        rhs = O.overwrite_rho(rho, mask, dval)
        if O.load().orc_have_pardiso():
            P = O.Pardiso(c.gx, c.gy, c.dx, c.dy, mask)
            phi_ref = P.solve(rhs)
            P.close()
        else:
            phi_ref = O.poisson_direct(c.gx, c.gy, c.dx, c.dy, mask, dval, rhs)
    phi = c.grid(hip.GRID_PHI)
    err = np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()
    print(f"{NX}x{NY} periodic={periodic} {solver}: {st['iterations']} PCG iterations, error {err:.2e}")
    assert err < 1e-10
    if solver == "direct":
        assert st["guess"] == 5 and st["iterations"] == 0, st
    else:
        assert st["iterations"] < 40, st
    c.close()


# the published voltages (200 V left, 0 V right and bottom) in the three
# rectangles a periodic geometry may hold (particles.cpp:651-654)
PUBLISHED_WALLS = [[0, 0, 0, 1, 200.0], [1, 0, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]


@pytest.mark.parametrize("case", ["per_64", "per_128x64_electrode", "published_500x256_electrode",
                                  "published_500x256_bottom_wall", "electrode_at_the_seam"])
def test_periodic_direct_solve(hip_mod, oracle_mod, case):
    """The direct solve with periodic y (DST-I in x, the real Fourier basis in
    y, capacitance correction for the internal Dirichlet nodes, transforms on
    the FP64 matrix-core GEMM) against the oracle's SuperLU solve of the
    reference's periodic matrix (generateCoeffsPeriodic,
    poissonSolver.cpp:82-180), to 1e-10 relative, with the PCG left nothing
    to do: the reference's published grid (501 x 257 nodes, periodic,
    Manya_Vansh_BMP.md:155-167) with an electrode and with its published
    boundary voltages (200 V left, 0 V right and bottom: row 0 held, so the
    last system row meets it through the wrap), and an electrode on rows 0-1
    whose neighbours lie across the periodic seam."""
    hip, O = hip_mod, oracle_mod
    dx = 1e-4
    NX, NY = 64, 64
    geom = PER_GEOM
    if case == "per_128x64_electrode":
        NX, NY = 128, 64
        geom = PER_GEOM[:2] + [[0.2, 0.1, 0.3, 0.3, -40.0]]
    elif case == "published_500x256_electrode":
        NX, NY = 500, 256
        geom = [[0, 0, 0, 1, 200.0], [1, 0, 1, 1, 0.0], [0.4375, 0.4375, 0.5625, 0.5625, 100.0]]
    elif case == "published_500x256_bottom_wall":
        NX, NY = 500, 256
        geom = PUBLISHED_WALLS
    elif case == "electrode_at_the_seam":
        NX, NY = 96, 128
        # rows 0 and 1: the row below it is the last system row (127)
        geom = PER_GEOM[:2] + [[0.25, 0.0, 0.5, 1 / 128, 30.0]]
    p = hip.default_params(NX=NX, NY=NY, width=NX * dx, height=NY * dx, dt=1.77e-11, layout=0, periodic=1,
                           poisson_tol=1e-12, poisson_direct=1)
    c = hip.Context(p)
    c.set_geometry(geom)
    rng = np.random.default_rng(NX + 3 * NY)
    for it in range(2):
        rho = rng.normal(0, 1e6, c.G)
        c.set_grid(hip.GRID_RHO, rho)
        st = c.poisson_solve(it)
        rhs, phi_ref, _, _ = O.poisson_periodic(rho, c.gx, c.gy, c.dx, c.dy, geom)
        phi = c.grid(hip.GRID_PHI)
        err = np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()
        print(f"{case} solve {it}: error {err:.2e}, relres {st['final_relres']:.1e}")
        assert err < 1e-10, (case, err, st)
        assert st["guess"] == 5 and st["iterations"] == 0, st
        assert st["final_relres"] <= 1e-12
        # the image row copies row 0 (its corners belong to the x walls, whose
        # potential may differ from a wall on row 0)
        p2 = phi.reshape(c.gy, c.gx)
        assert np.array_equal(p2[-1, 1:-1], p2[0, 1:-1])
        assert np.array_equal(c.grid(hip.GRID_RHO), rhs)
        ex, ey = O.efield_periodic(phi, c.gx, c.gy, c.dx, c.dy)
        assert np.array_equal(c.grid(hip.GRID_EX), ex) and np.array_equal(c.grid(hip.GRID_EY), ey)
    c.close()
