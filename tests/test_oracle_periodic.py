"""The periodic-y oracle pinned on the CPU: the reference's periodic operator
(generateCoeffsPeriodic, poissonSolver.cpp:82-180) against the discrete
eigenfunctions of the 5-point Laplacian with Dirichlet x walls and a periodic
y; its solve against cyclic-shift invariance; the periodic E field
(poissonSolver.cpp:411-430) against closed forms; the mover's modulo
(mover.cpp:13-22) against hand cases; the periodic push against a wrap of the
non-periodic push."""
import math

import numpy as np

WALLS_LR = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0]]


def _mask_lr(gx, gy):
    m = np.zeros(gx * gy, np.uint8)
    m.reshape(gy, gx)[:, 0] = 1
    m.reshape(gy, gx)[:, -1] = 1
    return m


def test_modulo_hand_cases(oracle_mod):
    O = oracle_mod
    assert O.modulo(1.25, 1.0) == 0.25
    assert O.modulo(-0.25, 1.0) == 0.75
    assert O.modulo(-2.0, 1.0) == 0.0           # integer quotient: 0
    assert O.modulo(0.0, 1.0) == 0.0
    # a tiny negative y wraps to exactly the height (then absorbed by y >= H)
    assert O.modulo(-1e-20, 0.0064) == 0.0064
    assert O.modulo(0.0064, 0.0064) == 0.0


def test_periodic_operator_eigenfunctions(oracle_mod):
    O = oracle_mod
    gx, gy, dx, dy = 17, 13, 1e-4, 2e-4
    P = gy - 1
    A = O.assemble_periodic(gx, gy, dx, dy, _mask_lr(gx, gy))
    assert A.shape == (gx * P, gx * P)
    i = np.arange(gx)[None, :]
    j = np.arange(P)[:, None]
    for k, m in [(1, 0), (2, 1), (5, 3), (16 - 1, 6)]:
        for phase in (np.cos, np.sin):
            u = (np.sin(math.pi * k * i / (gx - 1)) * phase(2 * math.pi * m * j / P)).ravel()
            lam = (2 - 2 * math.cos(math.pi * k / (gx - 1))) / dx**2 + (2 - 2 * math.cos(2 * math.pi * m / P)) / dy**2
            r = A @ u
            inner = np.ones((P, gx), bool)
            inner[:, [0, -1]] = False
            inner = inner.ravel()
            assert np.allclose(r[inner], lam * u[inner], rtol=0, atol=1e-9 * lam)
            assert np.allclose(r[~inner], -u[~inner])   # Dirichlet rows: -1 on the diagonal


def test_periodic_solve_is_shift_invariant_and_periodic(oracle_mod):
    O = oracle_mod
    gx, gy, dx, dy = 33, 33, 1e-4, 1e-4
    P = gy - 1
    rects = WALLS_LR + [[0.0, 0.0, 0.0, 1.0, 5.0]]      # left wall at 5 V (later rows win)
    rng = np.random.default_rng(3)
    rho = rng.normal(0, 1e5, gx * gy)
    rhs, phi, ex, ey = O.poisson_periodic(rho, gx, gy, dx, dy, rects)
    p2 = phi.reshape(gy, gx)
    assert np.array_equal(p2[-1], p2[0])                  # image row
    assert (p2[:, 0] == 5.0).all() and (p2[:, -1] == 0.0).all()
    # residual of the interior equations (wrap in y)
    lap = ((2 * p2[:P, 1:-1] - p2[:P, :-2] - p2[:P, 2:]) / dx**2
           + (2 * p2[:P, 1:-1] - np.roll(p2[:P], 1, 0)[:, 1:-1] - np.roll(p2[:P], -1, 0)[:, 1:-1]) / dy**2)
    r2 = rhs.reshape(gy, gx)[:P, 1:-1]
    assert np.abs(lap - r2).max() < 1e-9 * np.abs(r2).max()
    # shifting rho by s rows (cyclically over the P real rows) shifts phi
    s = 7
    rr = rho.reshape(gy, gx).copy()
    rr[:P] = np.roll(rr[:P], s, 0)
    rr[P] = rr[0]
    _, phi_s, _, _ = O.poisson_periodic(rr.ravel(), gx, gy, dx, dy, rects)
    want = np.roll(p2[:P], s, 0)
    assert np.abs(phi_s.reshape(gy, gx)[:P] - want).max() < 1e-9 * np.abs(want).max()


def test_periodic_efield_closed_form(oracle_mod):
    O = oracle_mod
    gx, gy, dx, dy = 21, 17, 1e-4, 1.5e-4
    P = gy - 1
    a, b = 300.0, 2.0
    jj = np.arange(gx)[None, :] * dx
    ii = np.arange(gy)[:, None]
    phi = a * jj + b * np.cos(2 * math.pi * (ii % P) / P) + 0 * jj
    ex, ey = O.efield_periodic(phi.ravel(), gx, gy, dx, dy)
    assert np.allclose(ex, -a, rtol=1e-9)
    want = b * np.sin(2 * math.pi * (ii % P) / P) * math.sin(2 * math.pi / P) / dy + 0 * jj
    assert np.allclose(ey.reshape(gy, gx), want, rtol=0, atol=1e-9 * np.abs(want).max())


def test_periodic_push_wraps_y(oracle_mod):
    O = oracle_mod
    NX = 32
    gx = gy = NX + 1
    dx = dy = 1e-4
    W = H = NX * dx
    dt = 1e-11
    G = gx * gy
    z = np.zeros(G)
    bz = np.zeros(G)
    # moving up through y = H, down through y = 0, and one staying inside
    x = np.array([0.5 * W, 0.5 * W, 0.3 * W])
    y = np.array([H - 0.2 * dy, 0.1 * dy, 0.5 * H])
    vy = np.array([0.5 * dy / dt, -0.4 * dy / dt, 0.1 * dy / dt])
    vx = np.zeros(3)
    vz = np.zeros(3)
    g = dict(gx=gx, gy=gy, dx=dx, dy=dy, width=W, height=H, qm=-1.0e11, dt=dt)
    X, Y, _, _, _, key, ab = O.push_f64(x, y, vx, vy, vz, z, z, bz, periodic=True, **g)
    assert (key >= 0).all() and ab.tolist() == [0, 0]
    y_free = y + dt * vy
    assert Y[0] == math.fmod(y_free[0], H) and Y[1] == H - math.fmod(-y_free[1], H) and Y[2] == y_free[2]
    # without the wrap the first two are absorbed
    _, _, _, _, _, key2, ab2 = O.push_f64(x, y, vx, vy, vz, z, z, bz, **g)
    assert key2.tolist()[:2] == [-1, -1] and ab2.tolist() == [0, 2]
    # CELL32: the same wrap in cell units
    cell, ox, oy, fvx, fvy, fvz = O.encode_cell32(x, y, vx, vy, vz, gx, dx, dy)
    rc, rox, roy, _, _, _, k3, ab3 = O.push_cell32(cell, ox, oy, fvx, fvy, fvz, z, z, bz, periodic=True, **g)
    assert (k3 >= 0).all() and ab3.tolist() == [0, 0]
    assert k3[0] // gx == 0 and k3[1] // gx == NX - 1
