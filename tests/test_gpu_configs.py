"""The build's workloads at their stated sizes on the GPU, against the
oracle (pic_configs.py names them; the physics is the one described there):

* plain_cpu_128 — 128^2 cells, 100k + 100k particles, Bz = 0.01 T, F64
  storage, 100 steps: every step against the oracle's own FP64 time step with
  MKL PARDISO (rho and phi to 1e-9, live counts and per-cell counts equal),
  and three steps checked stage by stage;
* electrode_512 — 512^2 cells, 4M + 4M, B tilted 30 deg, CELL32 storage: the
  first three steps stage by stage and against the oracle's FP64 time step
  with PARDISO (within FP32 storage rounding), 500 steps with property checks
  on every step, then one more step stage by stage;
* electrode_1024, precision form — 1024^2 cells, 2^24 + 2^24 particles,
  FP32 V-cycles with weighted-Jacobi smoothing (2 + 2 sweeps) inside FP64
  iterative refinement to 1e-10, 2^-32 fixed-point deposit: 1000 steps, then
  one step stage by stage with each field's error against the FP64 oracle and
  the deposit's error against the FP64 reference deposit, bounded by the
  fixed-point quantisation;
* electrode_1024, throughput form — the same grid and particles with the
  defaults the bench runs (2^-40 deposit, direct solve checked by the FP64
  PCG): 50 steps, then one step stage by stage, phi against PARDISO;
* the 64-run uneven ensemble (seed 4): three batched steps, four runs of
  different sizes checked stage by stage after every step.

"Stage by stage" is stage_check.check_step: deposit, rho, E and push bit for
bit against the oracle on the GPU's own inputs of each stage, phi against
PARDISO on the same right-hand side."""
import numpy as np
import pytest

import stage_check as SC

pytestmark = pytest.mark.gpu


def _workload(name):
    import pic_configs as cfg
    return cfg, cfg.WORKLOADS[name]


def _context(cfg, w, **kw):
    parts = [cfg.generate(w, s) for s in range(2)]
    return cfg.make_context(w, particles=parts, **kw), parts


def _sim(O, cfg, w, parts):
    sim = O.Sim(w.NX, w.NY, w.width, w.height, w.dt, cfg.EPS, w.bz, w.rects(), bx=w.bx)
    for s, (q, m, _) in enumerate(w.species()):
        sim.add_species(q, m, w.n_per_species, cfg.N0, *parts[s])
    return sim


def _relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _properties(hip, ctx, st, tol, full=False):
    """Cheap invariants of a step: the solve met its tolerance with finite
    residuals, the per-cell counts add up to the live count; with full=True
    also the stored state itself (cells sorted and inside the grid, offsets in
    [0, 1), finite velocities)."""
    assert np.isfinite(st["final_relres"]) and st["final_relres"] <= tol, st
    for s in range(2):
        n = ctx.count(s)
        counts = ctx.cell_counts(s)
        assert int(counts.sum()) == n
        if full:
            cell, ox, oy, vx, vy, vz = ctx.download_cell32(s)
            assert np.all(np.diff(cell) >= 0)
            gx = cell % ctx.gx
            gy = cell // ctx.gx
            assert np.all((gx < ctx.gx - 1) & (gy < ctx.gy - 1) & (cell >= 0))
            assert np.all((ox >= 0) & (ox < 1) & (oy >= 0) & (oy < 1))
            assert np.all(np.isfinite(vx) & np.isfinite(vy) & np.isfinite(vz))
            assert np.array_equal(np.bincount(cell, minlength=ctx.G), counts)
        assert np.isfinite(ctx.kinetic_energy(s))


@pytest.mark.timeout(300)
def test_plain_128_100_steps_vs_oracle(hip_mod, oracle_mod):
    hip, O = hip_mod, oracle_mod
    cfg, w = _workload("plain_cpu_128")
    assert (w.NX, w.n_per_species, w.layout, w.steps) == (128, 100_000, 0, 100)
    ctx, parts = _context(cfg, w, poisson_tol=1e-12)
    sim = _sim(O, cfg, w, parts)
    pref = SC.PoissonRef(O, ctx.gx, ctx.gy, ctx.dx, ctx.dy, w.rects())
    spec = SC.species_table(w, cfg.N0)
    worst = {"rho": 0.0, "phi": 0.0}
    try:
        for it in range(w.steps):
            staged = it in (0, 50, w.steps - 1)
            if staged:
                before, ab = SC.snapshot(ctx, 2)
            ctx.step(it)
            sim.step()
            if staged:
                errs = SC.check_step(hip, O, ctx, before, ab, spec, pref)
                assert errs["phi"] <= 1e-10, (it, errs)
            for name, which in (("rho", hip.GRID_RHO), ("phi", hip.GRID_PHI)):
                e = _relerr(ctx.grid(which), sim.grid(which))
                worst[name] = max(worst[name], e)
                assert e <= 1e-9, (it, name, e)
            for s in range(2):
                assert ctx.count(s) == sim.count(s), (it, s)
                if it % 10 == 9:
                    want = np.bincount(sim.particles(s)[5], minlength=ctx.G)
                    assert np.array_equal(ctx.cell_counts(s), want), (it, s)
        print(f"plain_cpu_128, 100 steps: worst rho {worst['rho']:.2e}, phi {worst['phi']:.2e}; "
              f"live {ctx.count(0)} + {ctx.count(1)}")
    finally:
        pref.close()
        sim.close()
        ctx.close()


@pytest.mark.timeout(400)
def test_electrode_512_500_steps(hip_mod, oracle_mod):
    hip, O = hip_mod, oracle_mod
    cfg, w = _workload("electrode_512")
    assert (w.NX, w.n_per_species, w.layout, w.steps, w.b_tilt_deg) == (512, 1 << 22, 1, 500, 30.0)
    tol = 1e-10
    ctx, parts = _context(cfg, w, poisson_tol=tol)
    sim = _sim(O, cfg, w, parts)
    pref = SC.PoissonRef(O, ctx.gx, ctx.gy, ctx.dx, ctx.dy, w.rects())
    spec = SC.species_table(w, cfg.N0)
    try:
        for it in range(3):
            before, ab = SC.snapshot(ctx, 2)
            st = ctx.step(it)
            sim.step()
            errs = SC.check_step(hip, O, ctx, before, ab, spec, pref)
            assert errs["phi"] <= 1e-10, (it, errs)
            # the oracle's own FP64 run from the same FP64 particles: the GPU
            # stores FP32 offsets and velocities, so rho differs by the
            # rounding of single particles' weights (2^-24 of a particle in a
            # node of ~64) and phi by far less
            er = _relerr(ctx.grid(hip.GRID_RHO), sim.grid(hip.GRID_RHO))
            ep = _relerr(ctx.grid(hip.GRID_PHI), sim.grid(hip.GRID_PHI))
            dn = [abs(ctx.count(s) - sim.count(s)) for s in range(2)]
            print(f"electrode_512 step {it}: stage errors {errs}; vs FP64 run rho {er:.2e} phi {ep:.2e} "
                  f"count differences {dn}")
            assert er <= 1e-6 and ep <= 1e-8, (it, er, ep)
            assert max(dn) <= 2, dn
            _properties(hip, ctx, st, tol)
        for it in range(3, w.steps):
            st = ctx.step(it)
            _properties(hip, ctx, st, tol, full=(it % 100 == 99))
        before, ab = SC.snapshot(ctx, 2)
        ctx.step(w.steps)
        errs = SC.check_step(hip, O, ctx, before, ab, spec, pref)
        print(f"electrode_512 step {w.steps}: {errs}; live {ctx.count(0)} + {ctx.count(1)}")
        assert errs["phi"] <= 1e-10, errs
    finally:
        pref.close()
        sim.close()
        ctx.close()


@pytest.mark.timeout(600)
def test_electrode_1024_precision_form_1000_steps(hip_mod, oracle_mod):
    hip, O = hip_mod, oracle_mod
    cfg, w = _workload("electrode_1024")
    assert (w.NX, w.n_per_species, w.layout, w.steps) == (1024, 1 << 24, 1, 1000)
    tol, cs = 1e-10, 32
    ctx, _ = _context(cfg, w, poisson_tol=tol, poisson_precision=1, mg_smoother=0, mg_pre=2, mg_post=2,
                      charge_shift=cs, poisson_direct=0)
    pref = SC.PoissonRef(O, ctx.gx, ctx.gy, ctx.dx, ctx.dy, w.rects())
    spec = SC.species_table(w, cfg.N0)
    try:
        its = []
        for it in range(w.steps):
            st = ctx.step(it)
            its.append((st["iterations"], st["inner_iterations"]))
            if it % 100 == 99 or it < 2:
                _properties(hip, ctx, st, tol, full=(it == w.steps - 1))
        before, ab = SC.snapshot(ctx, 2)
        st = ctx.step(w.steps)
        assert st["final_relres"] <= tol
        errs = SC.check_step(hip, O, ctx, before, ab, spec, pref, charge_shift=cs)
        # the deposit against the FP64 reference deposit: each of a particle's
        # four weights is rounded to 2^-32, so node n is within
        # (particles in its four cells) * 2^-33 of the exact sum (in weights)
        dep = []
        for s, (q, m, W) in enumerate(spec):
            cell, ox, oy, vx, vy, vz = before[s]
            xd = ((cell % ctx.gx) + ox.astype(np.float64)) * ctx.dx
            yd = ((cell // ctx.gx) + oy.astype(np.float64)) * ctx.dy
            fm_ref, _, _ = O.deposit_ref(xd, yd, vx, vy, vz, cell, ctx.gx, ctx.gy, ctx.dx, ctx.dy, W, q / m, q,
                                         cfg.EPS)
            fm = ctx.grid(hip.GRID_FINAL0 + s)
            cnt = np.bincount(cell, minlength=ctx.G).reshape(ctx.gy, ctx.gx)
            near = cnt.copy()
            near[:, 1:] += cnt[:, :-1]
            near[1:, :] += cnt[:-1, :]
            near[1:, 1:] += cnt[:-1, :-1]
            bound = near.reshape(-1) * np.ldexp(1.0, -cs - 1) * W * ctx.dx * ctx.dy * (1 + 1e-9) + 1e-9 * np.abs(
                fm_ref)
            assert np.all(np.abs(fm - fm_ref) <= bound), s
            dep.append(_relerr(fm, fm_ref))
        errs["deposit_vs_fp64"] = dep
        mean_it = np.mean(np.array(its, float), axis=0)
        print(f"electrode_1024 precision form, step {w.steps}: field errors vs the FP64 oracle {errs}; "
              f"refinement steps {mean_it[0]:.2f}, FP32 PCG iterations {mean_it[1]:.2f} per solve")
        assert errs["phi"] <= 1e-8 and errs["Ex"] <= 1e-6 and errs["Ey"] <= 1e-6, errs
    finally:
        pref.close()
        ctx.close()


@pytest.mark.timeout(600)
def test_electrode_1024_throughput_form_stage_by_stage(hip_mod, oracle_mod):
    """The throughput workload exactly as the bench steps it: electrode_1024
    (1024^2 cells, 2^24 + 2^24 particles, CELL32, 2^-40 deposit, B tilted 30
    deg), the direct solve checked by the FP64 PCG at 1e-10 (Chebyshev
    smoothing, previous-phi guess).  50 steps, every one leaving the PCG no
    iteration, with property checks; then one step stage by stage against the
    oracle (deposit, rho, E and push bit for bit) and phi against PARDISO."""
    hip, O = hip_mod, oracle_mod
    cfg, w = _workload("electrode_1024")
    assert (w.NX, w.n_per_species, w.layout) == (1024, 1 << 24, 1)
    tol, steps = 1e-10, 50
    ctx, _ = _context(cfg, w, poisson_tol=tol, poisson_guess=0, poisson_precision=0, mg_smoother=1,
                      capacity_factor=1.05)
    pref = SC.PoissonRef(O, ctx.gx, ctx.gy, ctx.dx, ctx.dy, w.rects())
    spec = SC.species_table(w, cfg.N0)
    try:
        for it in range(steps):
            st = ctx.step(it)
            assert st["iterations"] == 0, (it, st)
            if it % 10 == 9 or it < 2:
                _properties(hip, ctx, st, tol, full=(it == steps - 1))
        before, ab = SC.snapshot(ctx, 2)
        st = ctx.step(steps)
        assert st["iterations"] == 0 and st["final_relres"] <= tol, st
        errs = SC.check_step(hip, O, ctx, before, ab, spec, pref)
        print(f"electrode_1024 throughput form, step {steps}: field errors vs PARDISO {errs}")
        assert errs["phi"] <= 1e-10 and errs["Ex"] <= 1e-8 and errs["Ey"] <= 1e-8, errs
    finally:
        pref.close()
        ctx.close()


@pytest.mark.timeout(600)
def test_ensemble_64_runs_stage_by_stage(hip_mod, oracle_mod):
    hip, O = hip_mod, oracle_mod
    import pic_configs as cfg

    runs = cfg.ensemble(64, seed=4)
    assert len(runs) == 64 and {w.NX for w in runs} == {64, 128, 256}
    sizes = [w.n_per_species for w in runs]
    checked = [int(np.argmax(sizes)), int(np.argmin(sizes))]
    for g in (64, 128, 256):
        cand = [k for k, w in enumerate(runs) if w.NX == g and k not in checked]
        checked.append(cand[0])
    checked = sorted(set(checked))
    ctxs = []
    try:
        for w in runs:
            ctxs.append(_context(cfg, w)[0])
        prefs = {k: SC.PoissonRef(O, ctxs[k].gx, ctxs[k].gy, ctxs[k].dx, ctxs[k].dy, runs[k].rects()) for k in checked}
        ens = hip.Ensemble(ctxs)
        try:
            for it in range(3):
                snaps = {k: SC.snapshot(ctxs[k], 2) for k in checked}
                ens.step(1, graph=True)
                for k in checked:
                    errs = SC.check_step(hip, O, ctxs[k], *snaps[k], SC.species_table(runs[k], cfg.N0), prefs[k])
                    assert errs["phi"] <= 1e-10, (it, k, errs)
                    if it == 2:
                        print(f"run {k} ({runs[k].NX}^2, {runs[k].n_per_species} per species): {errs}")
        finally:
            ens.close()
            for p in prefs.values():
                p.close()
    finally:
        for c in ctxs:
            c.close()
