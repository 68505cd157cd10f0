"""Edge cases of the time step against the oracle, stage by stage (deposit,
rho, E and the push bit for bit, phi against PARDISO, absorbed counts), in
both storage layouts on 64 x 64 cells:

* empty_and_single: no electrons at all, one ion;
* all_leave: every electron sits within half a cell of the right wall moving
  out at 5 cells per step, so the first push absorbs all of them (none at the
  left wall); the later steps run with the species empty beside 5000 ions."""
import numpy as np
import pytest

import stage_check as SC

pytestmark = pytest.mark.gpu


def _particles(cfg, w, case):
    ions = cfg.generate(w, 1)
    if case == "empty_and_single":
        e = tuple(np.zeros(0) for _ in range(5))
        ions = tuple(a[:1] for a in ions)
        return [e, ions]
    rng = np.random.default_rng(5)
    n = 3000
    x = w.width - rng.uniform(0.01, 0.5, n) * w.dx
    y = rng.uniform(0.0, 1.0, n) * w.height
    vx = np.full(n, 5.0 * w.dx / w.dt)
    vy = rng.normal(0.0, 1e5, n)
    vz = rng.normal(0.0, 1e5, n)
    return [(x, y, vx, vy, vz), ions]


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("case", ["empty_and_single", "all_leave"])
def test_step_edge_cases_stage_by_stage(hip_mod, oracle_mod, layout, case):
    hip, O = hip_mod, oracle_mod
    import pic_configs as cfg

    w = cfg.WORKLOADS["plain_cpu_128"].scaled(NX=64, n=5000, name=f"edges_{case}")
    parts = _particles(cfg, w, case)
    ctx = cfg.make_context(w, layout=layout, particles=parts, poisson_tol=1e-12)
    pref = SC.PoissonRef(O, ctx.gx, ctx.gy, ctx.dx, ctx.dy, w.rects())
    spec = SC.species_table(w, cfg.N0)
    try:
        assert [ctx.count(s) for s in range(2)] == [len(parts[0][0]), len(parts[1][0])]
        for it in range(3):
            before, ab = SC.snapshot(ctx, 2)
            st = ctx.step(it)
            assert st["final_relres"] <= 1e-12, (it, st)
            errs = SC.check_step(hip, O, ctx, before, ab, spec, pref)
            assert errs["phi"] <= 1e-10, (it, errs)
            for s in range(2):
                assert int(ctx.cell_counts(s).sum()) == ctx.count(s), (it, s)
            if case == "all_leave" and it == 0:
                # every electron absorbed at the right wall in the first push
                assert ctx.count(0) == 0
                assert ctx.absorbed(0) == (0, len(parts[0][0]))
        if case == "empty_and_single":
            assert ctx.count(0) == 0 and ctx.absorbed(0) == (0, 0)
            assert ctx.count(1) + ctx.absorbed(1)[1] == 1
    finally:
        pref.close()
        ctx.close()
