"""Stage-by-stage check of one GPU time step against the oracle.

The oracle is applied to the GPU's own inputs of each stage, so every stage is
held to its own bar, whatever the length of the run before it:

* deposit (src/interpolation.cpp:36-171): the fixed-point node sums of every
  species equal the oracle's fixed-point deposit of the state the step started
  from, bit for bit; final_mesh is those sums times W dx dy 2^-shift, bit for
  bit; rho (after overwriteRho, src/poissonSolver.cpp:384-402) bit for bit;
* Poisson (src/poissonSolver.cpp:459-544): phi against MKL PARDISO (or the
  oracle's band Cholesky) on the same right-hand side;
* E (src/poissonSolver.cpp:433-456): bit for bit from the GPU's phi;
* push (src/mover.cpp:103-176,304-367): the oracle's push of the starting
  state in the GPU's fields equals the stored state after the step, as a set of
  particles (the order inside a cell follows the slot atomics), bit for bit;
  the absorption counters grow by the oracle's absorbed counts.

Test infrastructure only (imported by tests)."""
from __future__ import annotations

import numpy as np

_MIX = (0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9, 0x94D049BB133111EB, 0xD6E8FEB86659FD93, 0xA24BAED4963EE407,
        0x9FB21C651E98DF25)


def canonical_order(cell, fields):
    """Indices that sort particles by cell, then by a 64-bit hash of all their
    fields: two arrays holding the same particles in any order inside a cell
    come out row for row equal."""
    h = np.zeros(len(cell), np.uint64)
    for k, a in enumerate(fields):
        a = np.ascontiguousarray(a)
        u = a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64).astype(np.uint64)
        h = (h ^ u) * np.uint64(_MIX[k % len(_MIX)])
        h ^= h >> np.uint64(31)
    return np.lexsort((h, np.asarray(cell)))


def same_particles(a, b):
    """a, b: (cell, f1, f2, ...) tuples; True when they hold the same rows."""
    if len(a[0]) != len(b[0]):
        return False
    oa = canonical_order(a[0], a[1:])
    ob = canonical_order(b[0], b[1:])
    return all(np.array_equal(np.asarray(x)[oa], np.asarray(y)[ob]) for x, y in zip(a, b))


def snapshot(ctx, nspecies):
    """Stored state (raw: CELL32 cell/offsets/velocities, or F64 positions,
    velocities and cell) and absorption counters of every species."""
    lay = ctx.params.layout
    st = []
    for s in range(nspecies):
        if lay == 1:
            st.append(ctx.download_cell32(s))
        else:
            x, y, vx, vy, vz, cell = ctx.download(s)
            st.append((cell, x, y, vx, vy, vz))
    return st, [ctx.absorbed(s) for s in range(nspecies)]


class PoissonRef:
    """FP64 reference solve of the reference system for one geometry:
    MKL PARDISO when the oracle has it, else the band Cholesky."""

    def __init__(self, O, gx, gy, dx, dy, geom):
        self.O = O
        _, _, self.mask, self.dval = O.geometry(gx, gy, geom)
        self.gx, self.gy, self.dx, self.dy = gx, gy, dx, dy
        self.P = O.Pardiso(gx, gy, dx, dy, self.mask) if O.load().orc_have_pardiso() else None

    def solve(self, rhs):
        if self.P is not None:
            return self.P.solve(rhs)
        return self.O.poisson_direct(self.gx, self.gy, self.dx, self.dy, self.mask, self.dval, rhs)

    def close(self):
        if self.P is not None:
            self.P.close()
            self.P = None


def check_step(hip, O, ctx, before, absorbed_before, species, pref, *, charge_shift=40, energy_shift=32,
               energy=True):
    """Check the step ctx just took from `before` (see snapshot()).

    species: (charge, mass, W) per species; pref: PoissonRef of the context's
    geometry.  Returns the relative (max-norm) errors of phi, Ex and Ey against
    the FP64 reference fields of the same rho, and the absorbed counts."""
    p = ctx.params
    gx, gy, dx, dy = ctx.gx, ctx.gy, ctx.dx, ctx.dy
    lay = p.layout
    rho = None
    for s, (q, m, W) in enumerate(species):
        qc, qe = ctx.deposit_fixed(s)
        if lay == 1:
            cell, ox, oy, vx, vy, vz = before[s]
            rqc, rqe = O.deposit_fixed_cell32(cell, ox, oy, vx, vy, vz, gx, gy, charge_shift, energy_shift, q / m)
        else:
            cell, x, y, vx, vy, vz = before[s]
            rqc, rqe = O.deposit_fixed_f64(x, y, vx, vy, vz, cell, gx, gy, dx, dy, charge_shift, energy_shift, q / m)
        assert np.array_equal(qc, rqc), f"species {s}: fixed-point charge differs from the oracle"
        if energy:
            assert np.array_equal(qe, rqe), f"species {s}: fixed-point energy differs from the oracle"
        fm = rqc.astype(np.float64) * (W * (dx * dy) * np.ldexp(1.0, -charge_shift))
        assert np.array_equal(ctx.grid(hip.GRID_FINAL0 + s), fm), f"species {s}: final_mesh"
        term = fm * (-q / p.epsilon)
        rho = term if rho is None else rho + term
    rhs = O.overwrite_rho(rho, pref.mask, pref.dval)
    assert np.array_equal(ctx.grid(hip.GRID_RHO), rhs), "rho after overwriteRho"
    phi = ctx.grid(hip.GRID_PHI)
    phi_ref = pref.solve(rhs)
    ex, ey = O.efield(phi, gx, gy, dx, dy)
    assert np.array_equal(ctx.grid(hip.GRID_EX), ex), "Ex"
    assert np.array_equal(ctx.grid(hip.GRID_EY), ey), "Ey"
    exr, eyr = O.efield(phi_ref, gx, gy, dx, dy)
    errs = {
        "phi": float(np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()),
        "Ex": float(np.abs(ex - exr).max() / np.abs(exr).max()),
        "Ey": float(np.abs(ey - eyr).max() / np.abs(eyr).max()),
    }
    bz = ctx.grid(hip.GRID_BZ)
    g = dict(gx=gx, gy=gy, dx=dx, dy=dy, width=p.width, height=p.height, dt=p.dt, bx=p.bx, by=p.by,
             periodic=bool(p.periodic))
    absorbed = []
    for s, (q, m, W) in enumerate(species):
        if lay == 1:
            rc, rox, roy, rvx, rvy, rvz, key, ab = O.push_cell32(*before[s], ex, ey, bz, qm=q / m, **g)
            live = key >= 0
            want = (key[live], rox[live], roy[live], rvx[live], rvy[live], rvz[live])
            got = ctx.download_cell32(s)
        else:
            cell, x, y, vx, vy, vz = before[s]
            X, Y, VX, VY, VZ, key, ab = O.push_f64(x, y, vx, vy, vz, ex, ey, bz, qm=q / m, **g)
            live = key >= 0
            want = (key[live], X[live], Y[live], VX[live], VY[live], VZ[live])
            gxs, gys, gvx, gvy, gvz, gcell = ctx.download(s)
            got = (gcell, gxs, gys, gvx, gvy, gvz)
        assert ctx.count(s) == int(live.sum()), f"species {s}: live count"
        assert same_particles(got, want), f"species {s}: pushed state differs from the oracle's push"
        left0, tot0 = absorbed_before[s]
        left1, tot1 = ctx.absorbed(s)
        assert (left1 - left0, tot1 - tot0) == (int(ab[0]), int(ab[1])), f"species {s}: absorbed counts"
        absorbed.append((int(ab[0]), int(ab[1])))
    errs["absorbed"] = absorbed
    return errs


def species_table(w, n0):
    """(charge, mass, W) of each species of workload w (the weighting factor
    of pic_add_species: W = actual * area / (sim dx^2 dy^2))."""
    dx = w.width / w.NX
    dy = w.height / w.NY
    W = n0 * (w.width * w.height) / (w.n_per_species * dx * dy * dx * dy)
    return [(q, m, W) for q, m, _ in w.species()]
