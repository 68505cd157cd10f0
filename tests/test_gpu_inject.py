"""GPU injection (re-injected electrons and ionization pairs, after the push)
against the oracle: the state after a step with injection equals the
oracle's push of the state before it plus the oracle's injected particles
(bit-exact for the pushed particles; injected ones to 1e-12 relative, the
device and host log / sin / cos / asin differing by an ulp or two), the
counters match the exits, and a larger CELL32 run keeps its storage sorted
with deposits bit-identical to the oracle's deposit of the downloaded set."""
import numpy as np
import pytest
from scipy import stats

import inject_oracle as IO

pytestmark = pytest.mark.gpu

WALLS = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]
E, ME, MI = 1.602176634e-19, 9.1093837015e-31, 40 * 1.66053906660e-27
DX = 1e-4


def _ctx(hip, NX, layout, dt=1.77e-11):
    p = hip.default_params(NX=NX, NY=NX, width=NX * DX, height=NX * DX, dt=dt, b0=0.01, layout=layout)
    c = hip.Context(p)
    c.set_geometry(WALLS)
    return c


def _state(hip, c, n_e, n_i, seed, n_exit_e, n_exit_i):
    """Maxwellian particles plus particles that leave through x < 0 in one step."""
    out = []
    for s, (q, m, T, n, nx) in enumerate([(-E, ME, 2.0, n_e, n_exit_e), (E, MI, 0.03, n_i, n_exit_i)]):
        x, y, vx, vy, vz = hip.generate_particles(seed, s, n, c.params.width, c.params.height, c.dx, c.dy,
                                                  abs(q / m), T)
        x[:nx] = 0.3 * c.dx
        vx[:nx] = -50 * c.dx / c.params.dt
        out.append((x, y, vx, vy, vz))
    return out


def _canon(*cols):
    a = np.stack([np.asarray(v, np.float64) for v in cols], 1)
    return a[np.lexsort(a.T[::-1])]


def _close_sets(got, want, rtol=1e-12):
    """Rows of got and want as multisets; sorted by vx (continuous, distinct)."""
    assert got.shape == want.shape, (got.shape, want.shape)
    g = got[np.argsort(got[:, 3], kind="stable")]
    w = want[np.argsort(want[:, 3], kind="stable")]
    assert np.array_equal(g[:, 0], w[:, 0])   # cells
    scale = np.abs(w[:, 1:]).max(axis=0)
    err = np.abs(g[:, 1:] - w[:, 1:]).max(axis=0) / scale
    assert (err <= rtol).all(), err


@pytest.mark.parametrize("x_inject", [0.003, 0.024])
def test_step_with_injection_matches_oracle(hip_mod, oracle_mod, x_inject):
    hip, O = hip_mod, oracle_mod
    NX = 64
    c = _ctx(hip, NX, hip.PIC_LAYOUT_F64)
    seed = 7
    parts = _state(hip, c, 6000, 5000, 3, 700, 250)
    qms = [-E / ME, E / MI]
    for s, (q, m) in enumerate([(-E, ME), (E, MI)]):
        c.add_species(q, m, 6000, 1e16, 20000)
        c.upload(s, *parts[s])
    c.set_injection(x_inject=x_inject, te_reinject_eV=2.0, te_eV=2.0, ti_eV=0.03, seed=seed)
    target, attempts = c.injection_target()
    assert (target, attempts) == IO.pair_target(c.params.height, c.params.dt)
    g = dict(gx=c.gx, gy=c.gy, dx=c.dx, dy=c.dy, width=c.params.width, height=c.params.height)
    state = [c.download(s) for s in range(2)]          # (x, y, vx, vy, vz, cell)
    tot = {"re": 0, "np": 0, "drop": 0}
    for it in range(3):
        c.step(it)
        ex, ey, bz = c.grid(hip.GRID_EX), c.grid(hip.GRID_EY), c.grid(hip.GRID_BZ)
        pushed, left = [], []
        for s in range(2):
            x, y, vx, vy, vz, _ = state[s]
            X, Y, VX, VY, VZ, key, ab = O.push_f64(x, y, vx, vy, vz, ex, ey, bz, qm=qms[s], dt=c.params.dt, **g)
            live = key >= 0
            pushed.append([key[live], X[live], Y[live], VX[live], VY[live], VZ[live]])
            left.append(int(ab[0]))
        ie, ii = IO.injection(seed, it, left[0], left[1], width=c.params.width, height=c.params.height,
                              dt=c.params.dt, qm_e=qms[0], qm_i=qms[1], x_inject=x_inject, te_reinject=2.0,
                              te=2.0, ti=0.03)
        st = c.injection_stats()
        n_re = IO.reinject_count(*left) if 0 <= x_inject < c.params.width else 0
        dropped = IO.reinject_count(*left) - n_re
        tot["re"] += n_re
        tot["np"] += target
        tot["drop"] += dropped
        assert st == {"injections": it + 1, "reinjected": tot["re"], "pairs": tot["np"], "dropped": tot["drop"],
                      "last_electrons": n_re + target, "last_ions": target}, (it, st, left)
        assert ie[0].size == n_re + target and ii[0].size == target
        new_state = []
        for s, inj in enumerate((ie, ii)):
            x, y, vx, vy, vz, cell = c.download(s)
            want = np.concatenate([np.stack(pushed[s], 1),
                                   np.stack([O.keys(inj[0], inj[1], c.gx, c.dx, c.dy), *inj], 1)])
            _close_sets(np.stack([cell, x, y, vx, vy, vz], 1), want)
            # storage sorted by cell and described by the cell starts
            assert (np.diff(cell) >= 0).all()
            assert np.array_equal(c.cell_counts(s), np.bincount(cell, minlength=c.G))
            new_state.append((x, y, vx, vy, vz, cell))
        state = new_state
    c.close()


def test_standalone_inject_and_capacity_clamp(hip_mod, oracle_mod):
    """pic_inject (the drop-in's generate_new_particles) after the pushes, on
    CELL32 storage: new particles appended and re-binned; the ion storage is
    nearly full, so part of the pairs' ions are dropped and counted."""
    hip, O = hip_mod, oracle_mod
    c = _ctx(hip, 64, hip.PIC_LAYOUT_CELL32)
    parts = _state(hip, c, 5000, 4000, 5, 400, 100)
    cap_i = 4050
    for s, (q, m, cap) in enumerate([(-E, ME, 20000), (E, MI, cap_i)]):
        c.add_species(q, m, 5000, 1e16, cap)
        c.upload(s, *parts[s])
    seed = 11
    c.set_injection(x_inject=0.002, te_reinject_eV=2.0, seed=seed)
    target, _ = c.injection_target()
    for s in range(2):
        c.push(s, rebin=True)
    left = [c.absorbed(s)[0] for s in range(2)]
    n1 = [c.count(0), c.count(1)]
    before = c.download_cell32(0)
    c.inject()
    st = c.injection_stats()
    n_re = IO.reinject_count(*left)
    room = cap_i - n1[1]
    assert 0 < room < target
    assert st == {"injections": 1, "reinjected": n_re, "pairs": target, "dropped": target - room,
                  "last_electrons": n_re + target, "last_ions": room}
    assert [c.count(0), c.count(1)] == [n1[0] + n_re + target, cap_i]
    # electrons: the pushed ones unchanged plus the oracle's injected ones
    ie, _ = IO.injection(seed, 0, left[0], left[1], width=c.params.width, height=c.params.height,
                         dt=c.params.dt, qm_e=-E / ME, qm_i=E / MI, x_inject=0.002, te_reinject=2.0)
    enc = O.encode_cell32(*ie, c.gx, c.dx, c.dy)
    want = np.concatenate([np.stack(before, 1).astype(np.float64), np.stack(enc, 1).astype(np.float64)])
    got = c.download_cell32(0)
    _close_sets(np.stack(got, 1).astype(np.float64), want, rtol=1e-6)
    for s in range(2):
        cell = c.download_cell32(s)[0]
        assert (np.diff(cell) >= 0).all()
        assert np.array_equal(c.cell_counts(s), np.bincount(cell, minlength=c.G))
    c.close()


def test_pairs_from_empty_storage_distribution(hip_mod):
    """Start with no particles: after one step the ions are exactly this step's
    pairs; their x follows the cosine profile of generate_pairs, y is uniform,
    velocities are normal with variance |q/m| T."""
    hip = hip_mod
    NX = 128
    c = _ctx(hip, NX, hip.PIC_LAYOUT_CELL32, dt=4e-10)
    for q, m in [(-E, ME), (E, MI)]:
        c.add_species(q, m, 1, 1e16, 400000)
    c.set_injection(te_eV=3.0, ti_eV=0.05, seed=5)
    target, attempts = c.injection_target()
    assert target > 5000
    c.step(0)
    assert c.count(1) == target and c.count(0) == target
    xi, yi, vxi, vyi, vzi, _ = c.download(1)
    xe, ye, vxe, _, _, _ = c.download(0)
    # same positions for the two species (float offsets: to a float ulp of the cell)
    a = _canon(xi, yi)
    b = _canon(xe, ye)
    assert np.allclose(a, b, rtol=0, atol=1e-7 * DX)
    W, H = c.params.width, c.params.height
    F = lambda x: (np.sin(np.clip((x - 0.00625) * IO.PI_REF / 0.0075, -np.pi / 2, np.pi / 2)) + 1) / 2
    acc = F(W)          # attempts kept: x < width
    cdf = lambda x: np.minimum(F(x) / acc, 1.0)
    assert stats.kstest(xi, cdf).pvalue > 1e-3
    assert stats.kstest(yi / H, "uniform").pvalue > 1e-3
    sig_i = np.sqrt(E / MI * 0.05)
    for v in (vxi, vyi, vzi):
        assert stats.kstest(v / sig_i, "norm").pvalue > 1e-3
    assert stats.kstest(vxe / np.sqrt(E / ME * 3.0), "norm").pvalue > 1e-3
    c.close()


def test_cell32_run_with_injection_keeps_storage_and_deposit_exact(hip_mod, oracle_mod):
    hip, O = hip_mod, oracle_mod
    NX = 256
    c = _ctx(hip, NX, hip.PIC_LAYOUT_CELL32)
    parts = _state(hip, c, 400000, 400000, 9, 3000, 1000)
    for s, (q, m) in enumerate([(-E, ME), (E, MI)]):
        c.add_species(q, m, 400000, 1e16, 600000)
        c.upload(s, *parts[s])
    c.set_injection(x_inject=0.01, te_reinject_eV=2.0, seed=3)
    target, _ = c.injection_target()
    counts = [c.count(0), c.count(1)]
    left0 = [c.absorbed(s)[1] for s in range(2)]
    for it in range(4):
        c.step(it)
        st = c.injection_stats()
        tot = [c.absorbed(s)[1] for s in range(2)]
        for s in range(2):
            added = st["last_electrons"] if s == 0 else st["last_ions"]
            assert c.count(s) == counts[s] - (tot[s] - left0[s]) + added, (it, s)
            counts[s] = c.count(s)
        left0 = tot
        assert st["last_ions"] == target
    qms = [-E / ME, E / MI]
    for s in range(2):
        cell, ox, oy, vx, vy, vz = c.download_cell32(s)
        assert (np.diff(cell) >= 0).all()
        assert np.array_equal(c.cell_counts(s), np.bincount(cell, minlength=c.G))
        c.deposit(s)
        qc, qe = c.deposit_fixed(s)
        rc, re_ = O.deposit_fixed_cell32(cell, ox, oy, vx, vy, vz, c.gx, c.gy, c.params.charge_shift,
                                         c.params.energy_shift, qms[s])
        assert np.array_equal(qc, rc) and np.array_equal(qe, re_)
    c.close()
