"""Pins the oracle's Poisson stage (overwriteRho, the reference 5-point system,
direct solve, overwritePhi, E field) against the reference's PARDISO dumps and
against closed forms."""
import numpy as np
import pytest

from conftest import GOLDEN

WALLS = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]


@pytest.mark.parametrize("it", [0, 1, 2, 3, 4])
def test_direct_solve_matches_reference_pardiso_dumps(oracle_mod, it):
    """validation_dumps/iter000{it}: PARDISO phi for rho_input at 257^2, dx = 1/256."""
    O = oracle_mod
    rho = np.load(GOLDEN / f"iter{it:04d}_rho_input.npy")
    ref = np.load(GOLDEN / f"iter{it:04d}_phi_cpp.npy")
    rhs, phi, ex, ey = O.poisson_solver(rho, 257, 257, 1 / 256, 1 / 256, WALLS)
    assert np.abs(phi - ref).max() / np.abs(ref).max() < 2e-12
    # walls are Dirichlet 0 in both
    p2 = phi.reshape(257, 257)
    assert np.all(p2[0] == 0) and np.all(p2[-1] == 0) and np.all(p2[:, 0] == 0) and np.all(p2[:, -1] == 0)
    # overwriteRho: walls set to V (=0) then everything negated
    r2 = rhs.reshape(257, 257)
    assert np.array_equal(r2[1:-1, 1:-1], -rho.reshape(257, 257)[1:-1, 1:-1])


def test_pardiso_matches_dumps(oracle_mod):
    O = oracle_mod
    if not O.load().orc_have_pardiso():
        pytest.skip("oracle built without MKL")
    _, _, mask, dval = O.geometry(257, 257, WALLS)
    P = O.Pardiso(257, 257, 1 / 256, 1 / 256, mask)
    try:
        for it in range(5):
            rho = np.load(GOLDEN / f"iter{it:04d}_rho_input.npy")
            ref = np.load(GOLDEN / f"iter{it:04d}_phi_cpp.npy")
            phi = P.solve(O.overwrite_rho(rho, mask, dval))
            assert np.abs(phi - ref).max() / np.abs(ref).max() < 2e-12
    finally:
        P.close()


def test_discrete_eigenfunction(oracle_mod):
    """phi = sin(pi a i/NX) sin(pi b j/NY) is an eigenvector of the 5-point
    operator with Dirichlet walls: -Lap_h phi = lam phi with
    lam = (4/dx^2) sin^2(pi a/2NX) + (4/dy^2) sin^2(pi b/2NY)."""
    O = oracle_mod
    NX, NY = 40, 24
    dx, dy = 0.3, 0.2
    gx, gy = NX + 1, NY + 1
    a, b = 3, 2
    i = np.arange(gx)[None, :]
    j = np.arange(gy)[:, None]
    phi_true = np.sin(np.pi * a * i / NX) * np.sin(np.pi * b * j / NY)
    lam = 4 / dx**2 * np.sin(np.pi * a / (2 * NX)) ** 2 + 4 / dy**2 * np.sin(np.pi * b / (2 * NY)) ** 2
    # reference convention: Lap_h phi = rho_dep  ->  rho_dep = -lam phi
    rho = (-lam * phi_true).reshape(-1)
    _, phi, ex, ey = O.poisson_solver(rho, gx, gy, dx, dy, WALLS)
    assert np.abs(phi - phi_true.reshape(-1)).max() < 1e-12


def test_csr_layout_matches_generate_coeffs(oracle_mod):
    """Row structure of generateCoeffs (poissonSolver.cpp:204-236): 1-based CSR,
    Dirichlet rows = single -1, interior rows 5 entries at i-GX, i-1, i, i+1, i+GX."""
    O = oracle_mod
    gx = gy = 9
    dx, dy = 0.5, 0.25
    _, _, mask, _ = O.geometry(gx, gy, WALLS)
    ia, ja, a = O.assemble(gx, gy, dx, dy, mask)
    n_dir = int(mask.sum())
    assert ia[0] == 1 and ia[-1] - 1 == a.size == n_dir + 5 * (gx * gy - n_dir)
    i = 4 * gx + 4                                  # an interior node
    row = slice(ia[i] - 1, ia[i + 1] - 1)
    assert list(ja[row]) == [i - gx + 1, i, i + 1, i + 2, i + gx + 1]
    assert np.allclose(a[row], [-1 / dy**2, -1 / dx**2, 2 * (1 / dx**2 + 1 / dy**2), -1 / dx**2, -1 / dy**2])
    row0 = slice(ia[0] - 1, ia[1] - 1)
    assert list(ja[row0]) == [1] and list(a[row0]) == [-1]


def test_electrode_geometry_nodes(oracle_mod):
    """readGeometry truncation: (int)((GRID-1) * frac), inclusive rectangles."""
    O = oracle_mod
    rects = WALLS + [[0.4375, 0.4375, 0.5625, 0.5625, 100.0]]
    nodes, pot, mask, dval = O.geometry(513, 513, rects)
    assert list(nodes[-1]) == [224, 224, 288, 288] and pot[-1] == 100.0
    m = mask.reshape(513, 513)
    assert m[224:289, 224:289].all() and not m[223, 250] and not m[250, 289]
    assert dval.reshape(513, 513)[256, 256] == 100.0
    # fractions that are not exact in binary truncate downwards
    nodes, _, _, _ = O.geometry(11, 11, [[0.3, 0.3, 0.7, 0.7, 1.0]])
    assert list(nodes[0]) == [int(10 * 0.3), int(10 * 0.3), int(10 * 0.7), int(10 * 0.7)]


def test_efield_closed_forms(oracle_mod):
    """Central differences are exact on linear and quadratic phi; the edge
    copy gives every node the value of its clamped interior neighbour."""
    O = oracle_mod
    gx, gy, dx, dy = 17, 13, 0.1, 0.2
    x = (np.arange(gx) * dx)[None, :]
    y = (np.arange(gy) * dy)[:, None]
    phi = 3.0 * x - 2.0 * y + 0 * x * y
    ex, ey = O.efield(phi.reshape(-1), gx, gy, dx, dy)
    assert np.allclose(ex, -3.0, rtol=1e-12) and np.allclose(ey, 2.0, rtol=1e-12)
    phi = (x**2 + 0.5 * y**2) + 0 * x * y
    ex, ey = O.efield(phi.reshape(-1), gx, gy, dx, dy)
    ex2, ey2 = ex.reshape(gy, gx), ey.reshape(gy, gx)
    jj = np.clip(np.arange(gx), 1, gx - 2)
    ii = np.clip(np.arange(gy), 1, gy - 2)
    assert np.allclose(ex2, np.broadcast_to(-2 * (jj * dx)[None, :], (gy, gx)), rtol=1e-10)
    assert np.allclose(ey2, np.broadcast_to(-(ii * dy)[:, None], (gy, gx)), rtol=1e-10)
