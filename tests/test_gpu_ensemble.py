"""Batched ensemble of uneven runs (persistent kernels, one captured HIP graph
per step) against serial runs of the same contexts.  Runs whose Poisson solve
is a batched direct solve match their serial twins bit for bit on every step
(rho, phi, live and per-cell counts: the same integer deposit, the same
transform passes, the same push); runs the ensemble's one-workgroup PCG
solves match to 1e-9 (another PCG implementation, both at 1e-12) and are
checked stage by stage against the oracle on their own inputs.  Graph
replay is identical to direct launches; the autotuner's choice is valid;
mixed ensembles (two batched groups, a GEMM-transform grid and a PCG-only
run)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _runs():
    import pic_configs as cfg
    runs = cfg.ensemble(nruns=5, seed=9, grids=(64, 128), n_min=5e3, n_max=6e4)
    return runs


def _contexts(hip, runs):
    import pic_configs as cfg
    out = []
    for w in runs:
        parts = [cfg.generate(w, s) for s in range(2)]
        out.append(cfg.make_context(w, layout=1, particles=parts, poisson_tol=1e-12))
    return out


def test_ensemble_matches_serial_runs(hip_mod):
    hip = hip_mod
    runs = _runs()
    ens_ctx = _contexts(hip, runs)
    ser_ctx = _contexts(hip, runs)
    ens = hip.Ensemble(ens_ctx)
    for it in range(3):
        ens.step(1, graph=True)
        for c in ser_ctx:
            c.step(it)
        iters, rr = ens.poisson_stats()
        assert (rr[:, 1] <= 1e-12).all(), rr
        for k, (a, b) in enumerate(zip(ens_ctx, ser_ctx)):
            ra, rb = a.grid(hip.GRID_RHO), b.grid(hip.GRID_RHO)
            pa, pb = a.grid(hip.GRID_PHI), b.grid(hip.GRID_PHI)
            # same kernels for the particles, the same direct-solve passes
            # (batched over the runs of one shape): the same numbers
            assert np.array_equal(ra, rb), (it, k)
            assert np.array_equal(pa, pb), (it, k, np.abs(pa - pb).max())
            for s in range(2):
                assert a.count(s) == b.count(s), (it, k, s)
                assert np.array_equal(a.cell_counts(s), b.cell_counts(s)), (it, k, s)
    # the runs stay usable through the ordinary API: one more serial step each
    for c in ens_ctx:
        c.step(3)
    ens.close()
    for c in ens_ctx + ser_ctx:
        c.close()


def test_ensemble_from_runs_stepped_serially(hip_mod):
    """Runs that have taken serial steps (their cell arrays left unwritten by
    the single-run step) form an ensemble: its first step equals a further
    serial step of twin runs (rho bit-identical, phi to 1e-9, counts within 2)."""
    hip = hip_mod
    runs = _runs()[:3]
    ens_ctx = _contexts(hip, runs)
    ser_ctx = _contexts(hip, runs)
    for it in range(2):
        for c in ens_ctx + ser_ctx:
            c.step(it)
    ens = hip.Ensemble(ens_ctx)
    ens.step(1, graph=True)
    for c in ser_ctx:
        c.step(2)
    for k, (a, b) in enumerate(zip(ens_ctx, ser_ctx)):
        assert np.array_equal(a.grid(hip.GRID_RHO), b.grid(hip.GRID_RHO)), k
        assert np.array_equal(a.grid(hip.GRID_PHI), b.grid(hip.GRID_PHI)), k
        for s in range(2):
            assert a.count(s) == b.count(s), (k, s)
    ens.close()
    for c in ens_ctx + ser_ctx:
        c.close()


def test_ensemble_graph_equals_direct_launches_and_tuner(hip_mod):
    hip = hip_mod
    runs = _runs()
    a_ctx = _contexts(hip, runs)
    b_ctx = _contexts(hip, runs)
    ea, eb = hip.Ensemble(a_ctx), hip.Ensemble(b_ctx)
    ea.step(3, graph=True)
    eb.step(3, graph=False)
    for a, b in zip(a_ctx, b_ctx):
        assert np.array_equal(a.grid(hip.GRID_PHI), b.grid(hip.GRID_PHI))
        for s in range(2):
            assert a.count(s) == b.count(s)
            assert np.array_equal(a.cell_counts(s), b.cell_counts(s))
    t = ea.tune(1)
    print("autotuner:", t)
    assert t["tile"] in (64, 128, 256) and t["blocks"] > 0 and t["ms_per_step"] > 0
    ea.step(2)
    it, rr = ea.poisson_stats()
    assert (rr[:, 1] <= 1e-12).all()
    ea.close()
    eb.close()
    for c in a_ctx + b_ctx:
        c.close()


def test_ensemble_mixed_direct_batches_and_pcg_runs(hip_mod, oracle_mod):
    """Runs of one ensemble solved different ways: batched direct solves in two
    groups of the same grid (the default electrode and a second geometry), a
    96-cell grid (sine transforms by GEMM: not batched, the ensemble's PCG
    solves it) and a run with the direct solve turned off.  Every run against
    its own serial steps, as above; the batched runs need no PCG iteration.
    The two PCG-solved runs are also checked stage by stage against the oracle
    on their own inputs (deposit, rho, E, push and absorbed counts bit for bit,
    phi against PARDISO): where their counts differ from the serial twins', the
    only cause is phi's rounding."""
    import pic_configs as cfg
    import stage_check as SC
    hip, O = hip_mod, oracle_mod
    runs = cfg.ensemble(nruns=6, seed=12, grids=(64, 96), n_min=5e3, n_max=4e4)
    runs[0] = cfg.Workload("mix0", 64, 64, 20000, 12000, 0.03, 0.0, cfg.ELECTRODE, electrode_v=50.0, layout=1)
    runs[1] = cfg.Workload("mix1", 64, 64, 15000, 12001, 0.02, 0.0, cfg.ELECTRODE, electrode_v=80.0, layout=1)
    runs[2] = cfg.Workload("mix2", 64, 64, 12000, 12002, 0.01, 0.0, cfg.ELECTRODE, electrode_v=120.0, layout=1)
    runs[3] = cfg.Workload("mix3", 96, 96, 12000, 12003, 0.01, 0.0, cfg.ELECTRODE, electrode_v=60.0, layout=1)
    other = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0],
             [0.2, 0.6, 0.3, 0.8, -40.0]]

    def build():
        out = []
        for k, w in enumerate(runs):
            parts = [cfg.generate(w, s) for s in range(2)]
            kw = {"poisson_direct": 0} if k == 2 else {}
            c = cfg.make_context(w, layout=1, particles=parts, poisson_tol=1e-12, **kw)
            if k == 1:
                c.set_geometry(other)
            out.append(c)
        return out

    ens_ctx, ser_ctx = build(), build()
    ens = hip.Ensemble(ens_ctx)
    pcg_runs = (2, 3)
    prefs = {k: SC.PoissonRef(O, ens_ctx[k].gx, ens_ctx[k].gy, ens_ctx[k].dx, ens_ctx[k].dy, runs[k].rects())
             for k in pcg_runs}
    for it in range(2):
        snaps = {k: SC.snapshot(ens_ctx[k], 2) for k in pcg_runs}
        ens.step(1, graph=True)
        for k in pcg_runs:
            errs = SC.check_step(hip, O, ens_ctx[k], *snaps[k], SC.species_table(runs[k], cfg.N0), prefs[k])
            assert errs["phi"] <= 1e-10, (it, k, errs)
        for c in ser_ctx:
            c.step(it)
        iters, rr = ens.poisson_stats()
        assert (rr[:, 1] <= 1e-12).all(), rr
        assert iters[0] == 0 and iters[1] == 0, iters        # batched direct solves
        if it == 0:
            assert iters[2] > 0, iters                        # PCG from zero
        for k, (a, b) in enumerate(zip(ens_ctx, ser_ctx)):
            pa, pb = a.grid(hip.GRID_PHI), b.grid(hip.GRID_PHI)
            if it == 0 or k in (0, 1, 4, 5):
                assert np.array_equal(a.grid(hip.GRID_RHO), b.grid(hip.GRID_RHO)), (it, k)
            if k in (0, 1, 4, 5):
                # batched direct solves: the single-run passes, the same numbers
                assert np.array_equal(pa, pb), (it, k)
                for s in range(2):
                    assert a.count(s) == b.count(s), (it, k, s)
            else:
                # solved by the ensemble's one-workgroup PCG against the serial
                # multi-kernel PCG, both to 1e-12: phi agrees to rounding, and
                # a particle within that rounding of a cell face or a wall may
                # end on the other side, so the counts may differ by a few
                assert np.abs(pa - pb).max() <= 1e-9 * np.abs(pb).max(), (it, k)
                for s in range(2):
                    assert abs(a.count(s) - b.count(s)) <= 2, (it, k, s)
    for pr in prefs.values():
        pr.close()
    ens.close()
    for c in ens_ctx + ser_ctx:
        c.close()
