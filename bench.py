"""Benchmark of the PIC time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload electrode_1024]

A step is one full time step (deposit both species, Poisson + E, push and
re-bin both species) over the workload's particles, resident in HBM.  Prints
one JSON line: particle-steps/s (sum over steps and species of live particles
divided by wall time), per-stage device times, the dominant kernel's roofline
and the CPU baseline (the oracle's restatement of the reference path, timed
on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "bmp-6th-sem_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes of the particle kernels, per particle (and per cell),
# for the CELL32 (1) and F64 (0) layouts.  CELL32 storage holds 2 offsets and
# 3 velocities (20 B); a particle's cell comes from the cell starts, so no
# kernel of the step reads or writes the cell array.  F64: x, y and 3
# velocities (40 B) plus the cell array, which its scatter writes.
#   k_push_tile    reads the state, writes the slot; per cell it reads the
#                  start and copies it to start_old;
#   k_scatter_push reads the state and the slot, writes the state to its sorted
#                  slot; per cell it reads the old start and the new start;
#   k_deposit_tile reads offsets (+ velocities for the energy mesh) per particle
#                  and, per cell, its start index, and writes two charge + two
#                  energy node-row sums (int64); the particles of the cell left
#                  of each tile, read a second time, are 1/256 of the whole and
#                  not counted.
KERNEL_BYTES = {
    "push": {1: (20 + 4, 4 + 4), 0: (40 + 4, 0)},
    "scatter": {1: (20 + 4 + 20, 4 + 4), 0: (40 + 4 + 44, 4)},
    "deposit": {1: (8 + 12, 4 + 32), 0: (16 + 24, 4 + 32)},
}
KERNEL_NAMES = {"push": "k_push_tile", "scatter": "k_scatter_push", "deposit": "k_deposit_tile"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB through the HIP runtime's
    copy kernel: the known-size reference that scales the HBM byte counters
    (FETCH_SIZE, WRITE_SIZE) read for the PIC kernels."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3):          # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))


def unet_only(n, reps=5):
    """Forward and forward + backward times of the network on an n x n image
    with seeded inputs, one training step first (its Adam update included)."""
    import numpy as np
    import pic_hip

    net = pic_hip.UNet(n, seed=1)
    rng = np.random.default_rng(0)
    x = rng.normal(0, 1, (n, n, 2)).astype(np.float32)
    net.forward(x)
    net.train_step(x, rng.normal(0, 0.1, (n, n)).astype(np.float32), np.ones((n, n), np.uint8), 1e-3)
    fwd, full = np.zeros(1, np.float32), np.zeros(1, np.float32)
    rc = pic_hip.load().pic_unet_benchmark(net.h, reps, pic_hip._fp(fwd), pic_hip._fp(full))
    net.close()
    fl = unet_flops(n)
    print(json.dumps({"unet_image": n, "rc": rc, "forward_ms": float(fwd[0]), "train_ms": float(full[0]),
                      "forward_gflop": fl / 1e9, "forward_tflops": fl / (float(fwd[0]) * 1e-3) / 1e12}))


def unet_flops(n):
    """Multiply-add FLOPs of one forward pass of the learned-guess network on an
    n x n image (the first layer's 2 input channels, not its 16-channel
    padding: the forward pass computes that layer with 2)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import unet_oracle as U
    f = 0
    for (src, _, _, cin_real, _, _, cout) in U.CONVS:
        lvl = U.TENSORS[src][0]
        f += 2 * 9 * cin_real * cout * (n >> lvl) ** 2
    return f + 2 * 16 * n * n


def learned_vs_plain(w, parts, layout, train, total, tol=1e-8):
    """The learned initial guess against the plain previous-solution warm start
    at relative residual `tol`: two contexts step the same particles for
    `total` steps, the learned one training online on the first `train`
    solves (its guess is offered from then on); Poisson ms, PCG iterations,
    starting residuals and the fallback rate are averaged over the solves
    after training.  Both use the PCG alone (poisson_direct = 0)."""
    import numpy as np
    import pic_configs as cfg

    plain = cfg.make_context(w, layout=layout, particles=parts, poisson_tol=tol, poisson_direct=0)
    learn = cfg.make_context(w, layout=layout, particles=parts, poisson_tol=tol, poisson_direct=0)
    learn.learned_enable(train_steps=train, min_trained=train, lr=1e-3, seed=1)
    t0 = time.perf_counter()
    sp, sl = [], []
    for it in range(total):
        # each context's step ends on the device before the other's begins
        # (a step returns with its push and scatter still running on its
        # stream): neither solve shares the GPU with the other's particles
        a = plain.step(it)
        plain.sync()
        b = learn.step(it)
        learn.sync()
        if it > train:
            sp.append(a)
            sl.append(b)
    elapsed = time.perf_counter() - t0
    st = learn.learned_stats()
    fwd_ms, train_ms = learn.learned_benchmark(5)
    n = (min(w.NX, w.NY) // 128) * 128
    fl = unet_flops(n)
    mean = lambda rows, k: float(np.mean([r[k] for r in rows]))
    out = {
        "grid": f"{w.NX}x{w.NY}", "particles_per_species": w.n_per_species, "image": n, "tol": tol,
        "steps": total, "train_steps": train, "eval_steps": len(sl),
        "poisson_ms_plain": mean(sp, "ms"), "poisson_ms_learned": mean(sl, "ms"),
        "iterations_plain": mean(sp, "iterations"), "iterations_learned": mean(sl, "iterations"),
        "initial_relres_plain": mean(sp, "initial_relres"), "initial_relres_learned": mean(sl, "initial_relres"),
        "guess_used": st["used"], "guess_tried": st["tried"],
        "fallback_rate": (1 - st["used"] / st["tried"]) if st["tried"] else None,
        "last_train_loss": st["loss"],
        "unet_forward_ms": fwd_ms, "unet_train_ms": train_ms,
        "unet_forward_gflop": fl / 1e9,
        "unet_mfma_tflops": fl / (fwd_ms * 1e-3) / 1e12 if fwd_ms > 0 else None,
        "unet_mfma_util": (fl / (fwd_ms * 1e-3)) / 2.5e15 if fwd_ms > 0 else None,
        "wall_s": elapsed,
    }
    out["speedup"] = out["poisson_ms_plain"] / out["poisson_ms_learned"] if out["poisson_ms_learned"] else None
    # what a better guess can buy: the plain PCG contracts the residual by rho
    # per iteration, so a guess that starts g times lower saves log g / log(1/rho)
    # iterations; at most every iteration's time can go
    ip, r0p, r0l = out["iterations_plain"], out["initial_relres_plain"], out["initial_relres_learned"]
    if ip > 0 and r0p > 0 and r0l > 0:
        rho = (tol / r0p) ** (1.0 / ip)
        ms_it = out["poisson_ms_plain"] / ip
        saved = np.log(r0p / r0l) / np.log(1.0 / rho)
        out["pcg_contraction_per_iteration"] = rho
        out["guess_gain"] = r0p / r0l
        out["iterations_saved_expected"] = saved
        out["speedup_bound_free_forward"] = ip / max(ip - saved, 1e-9)
        out["guess_gain_for_1p5x_free_forward"] = float((1.0 / rho) ** (ip / 3.0))
        out["ms_per_pcg_iteration"] = ms_it
    plain.close()
    learn.close()
    log(f"learned vs plain @ {tol:g} on {w.NX}^2: {out['poisson_ms_learned']:.3f} vs {out['poisson_ms_plain']:.3f} ms, "
        f"{out['iterations_learned']:.2f} vs {out['iterations_plain']:.2f} iterations, used {st['used']}/{st['tried']}, "
        f"U-Net forward {fwd_ms:.3f} ms")
    return out


def precision_form(steps, warmup=3):
    """The precision form at its configuration (electrode_1024, 2^24 + 2^24
    particles, 2^-32 fixed-point deposit, FP32 V-cycles with weighted-Jacobi
    2 + 2 sweeps inside FP64 iterative refinement to 1e-10): ms per step and
    the Poisson statistics; then, on the last step's right-hand side, each
    field's error against the library's FP64 direct solve of the same system
    (checked against MKL PARDISO in the tests)."""
    import numpy as np
    import pic_configs as cfg
    import pic_hip

    w = cfg.WORKLOADS["electrode_1024"]
    parts = [cfg.generate(w, s) for s in range(2)]
    ctx = cfg.make_context(w, particles=parts, poisson_tol=1e-10, poisson_precision=1, mg_smoother=0, mg_pre=2,
                           mg_post=2, charge_shift=32, poisson_direct=0)
    for it in range(warmup):
        ctx.step(it)
    ctx.sync()
    n0 = sum(ctx.count(s) for s in range(2))
    st = []
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        st.append(ctx.step(it))
    ctx.sync()
    el = time.perf_counter() - t0
    n1 = sum(ctx.count(s) for s in range(2))
    rhs = ctx.grid(pic_hip.GRID_RHO)
    fields = {k: ctx.grid(v) for k, v in (("phi", pic_hip.GRID_PHI), ("Ex", pic_hip.GRID_EX), ("Ey", pic_hip.GRID_EY))}
    ref = pic_hip.Context(pic_hip.default_params(NX=w.NX, NY=w.NY, width=w.width, height=w.height, dt=w.dt,
                                                 epsilon=cfg.EPS, layout=1, poisson_tol=1e-12))
    ref.set_geometry(w.rects())
    ref.set_grid(pic_hip.GRID_RHO, -rhs)     # overwriteRho of -rhs gives rhs again
    ref.poisson_solve(0)
    errs = {}
    for k, v in (("phi", pic_hip.GRID_PHI), ("Ex", pic_hip.GRID_EX), ("Ey", pic_hip.GRID_EY)):
        r = ref.grid(v)
        errs[k] = float(np.abs(fields[k] - r).max() / np.abs(r).max())
    ref.close()
    ctx.close()
    out = {"workload": w.name, "steps": steps, "ms_per_step": 1e3 * el / steps,
           "particle_steps_per_s": steps * 0.5 * (n0 + n1) / el,
           "poisson_ms": float(np.mean([s["ms"] for s in st])),
           "refinement_steps": float(np.mean([s["iterations"] for s in st])),
           "fp32_pcg_iterations": float(np.mean([s["inner_iterations"] for s in st])),
           "final_relres_max": float(max(s["final_relres"] for s in st)),
           "charge_shift": 32, "smoother": "weighted Jacobi 2+2",
           "field_errors_vs_fp64_direct": errs}
    log(f"precision form: {out['ms_per_step']:.3f} ms/step, Poisson {out['poisson_ms']:.3f} ms, field errors {errs}")
    return out


# The reference's published times for its periodic 501 x 257 run, 10 steps
# (Manya_Vansh_BMP.md:396-406: RTX 6000 Ada, PARDISO on the host), in ms/step
PUBLISHED_MS = {"deposit": 54.4, "mover": 34.2, "poisson": 78.1, "total": 213.0}


def published_workload(steps, warmup=3):
    """The reference's own published run (periodic_501x257: 501 x 257 nodes,
    periodic y, 9.6M electrons + 9.6M Xe+ ions) stepped on the GPU, once with
    the direct solve (DST in x, Fourier basis in y, matrix-core GEMMs) and
    once with the PCG alone: ms per step, particle-steps/s and the per-stage
    device times, beside the reference's published per-step times for the
    same workload on its own hardware."""
    import numpy as np
    import pic_configs as cfg

    w = cfg.WORKLOADS["periodic_501x257"]
    parts = [cfg.generate(w, s) for s in range(2)]
    out = {"workload": w.name, "cells": f"{w.NX}x{w.NY}", "nodes": f"{w.NX + 1}x{w.NY + 1}", "periodic_y": True,
           "particles_per_species": w.n_per_species, "steps": steps,
           "published_ms_per_step": PUBLISHED_MS,
           "published_hardware": "NVIDIA RTX 6000 Ada (deposit, mover) + host MKL PARDISO (Poisson)"}
    for solver, direct in (("direct", 1), ("pcg", 0)):
        ctx = cfg.make_context(w, particles=parts, poisson_tol=1e-10, poisson_direct=direct)
        it = 0
        for _ in range(warmup):
            ctx.step(it)
            it += 1
        ctx.sync()
        n0 = sum(ctx.count(s) for s in range(2))
        st = []
        t0 = time.perf_counter()
        for _ in range(steps):
            st.append(ctx.step(it))
            it += 1
        ctx.sync()
        el = time.perf_counter() - t0
        n1 = sum(ctx.count(s) for s in range(2))
        nb = min(10, steps)
        ctx.set_timing(1)
        ctx.stage_times()
        for _ in range(nb):
            ctx.step(it)
            it += 1
        stage = {k: v[0] / nb for k, v in ctx.stage_times().items()}
        ctx.set_timing(0)
        ctx.close()
        ms = 1e3 * el / steps
        r = {"ms_per_step": ms, "particle_steps_per_s": steps * 0.5 * (n0 + n1) / el,
             "stages_ms_per_step": {k: round(v, 4) for k, v in stage.items()},
             "poisson_iterations_mean": float(np.mean([x["iterations"] for x in st])),
             "poisson_final_relres_max": float(max(x["final_relres"] for x in st))}
        if solver == "direct":
            dep = stage["deposit"] + stage["finalize"]
            mover = stage["push"] + stage["bin"] + stage["scatter"]
            r["published_over_measured"] = {
                "deposit": PUBLISHED_MS["deposit"] / dep if dep > 0 else None,
                "mover": PUBLISHED_MS["mover"] / mover if mover > 0 else None,
                "poisson": PUBLISHED_MS["poisson"] / stage["poisson"] if stage["poisson"] > 0 else None,
                "total": PUBLISHED_MS["total"] / ms}
        out[solver] = r
        log(f"published workload, {solver} solve: {ms:.3f} ms/step, Poisson {stage['poisson']:.3f} ms "
            f"({r['poisson_iterations_mean']:.1f} PCG iterations)")
    return out


def ensemble_form(nruns, steps, seed=4):
    """Aggregate particle-steps/s of an uneven ensemble stepped as one batch
    (persistent kernels, one HIP graph per step, autotuned) against the same
    runs stepped one after another with the single-run path; plus a check of
    the first runs against serial copies after one step."""
    import numpy as np
    import pic_configs as cfg
    import pic_hip

    runs = cfg.ensemble(nruns=nruns, seed=seed)
    t0 = time.perf_counter()
    parts = [[cfg.generate(w, s) for s in range(2)] for w in runs]
    log(f"ensemble: {nruns} runs, {sum(2 * w.n_per_species for w in runs)} particles generated in "
        f"{time.perf_counter() - t0:.1f} s")

    def contexts():
        return [cfg.make_context(w, layout=1, particles=p, poisson_tol=1e-10) for w, p in zip(runs, parts)]

    def live(ctxs):
        return sum(c.count(s) for c in ctxs for s in range(2))

    # serial: each run stepped by the single-run path
    ser = contexts()
    for k, c in enumerate(ser):
        c.step(0)
    for c in ser:
        c.sync()
    n0 = live(ser)
    t1 = time.perf_counter()
    for it in range(1, 1 + steps):
        for c in ser:
            c.step(it)
    for c in ser:
        c.sync()
    t_ser = time.perf_counter() - t1
    rate_ser = steps * 0.5 * (n0 + live(ser)) / t_ser
    # batched
    ens_ctx = contexts()
    ens = pic_hip.Ensemble(ens_ctx)
    ens.step(1)
    tuned = ens.tune(2)
    n0 = live(ens_ctx)
    t1 = time.perf_counter()
    ens.step(steps)
    t_ens = time.perf_counter() - t1
    rate_ens = steps * 0.5 * (n0 + live(ens_ctx)) / t_ens
    it, rr = ens.poisson_stats()
    ens.close()
    for c in ens_ctx + ser:
        c.close()
    # per-run check: a fresh ensemble of the first runs against serial copies, one step
    k = min(8, nruns)
    a = [cfg.make_context(w, layout=1, particles=p, poisson_tol=1e-12) for w, p in zip(runs[:k], parts[:k])]
    b = [cfg.make_context(w, layout=1, particles=p, poisson_tol=1e-12) for w, p in zip(runs[:k], parts[:k])]
    e2 = pic_hip.Ensemble(a)
    e2.step(1)
    for c in b:
        c.step(0)
    rho_exact = all(np.array_equal(x.grid(pic_hip.GRID_RHO), y.grid(pic_hip.GRID_RHO)) for x, y in zip(a, b))
    phi_err = max(float(np.abs(x.grid(pic_hip.GRID_PHI) - y.grid(pic_hip.GRID_PHI)).max()
                        / np.abs(y.grid(pic_hip.GRID_PHI)).max()) for x, y in zip(a, b))
    e2.close()
    for c in a + b:
        c.close()
    out = {
        "runs": nruns, "steps": steps, "grids": sorted({w.NX for w in runs}),
        "particles": int(sum(2 * w.n_per_species for w in runs)),
        "value": rate_ens, "unit": "particle-steps/s",
        "serial_value": rate_ser, "speedup_vs_serial": rate_ens / rate_ser,
        "ms_per_step": 1e3 * t_ens / steps, "serial_ms_per_step": 1e3 * t_ser / steps,
        "autotuner": tuned, "poisson_iterations_max": int(it.max()), "poisson_final_relres_max": float(rr[:, 1].max()),
        "check_runs": k, "check_rho_bit_identical": bool(rho_exact), "check_phi_max_rel_err": phi_err,
    }
    log(f"ensemble: {rate_ens:.3e} particle-steps/s batched vs {rate_ser:.3e} serial "
        f"({out['speedup_vs_serial']:.1f}x), tuned {tuned}")
    return out


def injection_form(ctx, w, it, steps):
    """The step with the reference's particle injection after the push
    (main.cpp:232-245): re-injected electrons for the x < 0 exits and the
    ionization pairs of generate_pairs, at the workload's temperatures."""
    (_, _, te), (_, _, ti) = w.species()
    cfg = ctx.set_injection(te_reinject_eV=float(int(te)), te_eV=te, ti_eV=ti, seed=7)
    target, _ = ctx.injection_target()
    ctx.step(it)                                  # warm-up (allocations settle)
    it += 1
    ctx.sync()
    n0 = sum(ctx.count(s) for s in range(2))
    st0 = ctx.injection_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step(it)
        it += 1
    ctx.sync()
    t1 = time.perf_counter()
    n1 = sum(ctx.count(s) for s in range(2))
    st1 = ctx.injection_stats()
    ctx.set_injection(False)
    ms = 1e3 * (t1 - t0) / steps
    return {"steps": steps, "ms_per_step": ms,
            "particle_steps_per_s": steps * 0.5 * (n0 + n1) / (t1 - t0),
            "pairs_per_step": target, "x_inject_m": cfg["x_inject"],
            "reinjected_per_step": (st1["reinjected"] - st0["reinjected"]) / steps,
            "dropped": st1["dropped"] - st0["dropped"]}, it


def cpu_baseline(w, parts, steps=3):
    """Oracle (C, OpenMP, FP64 AoS, MKL PARDISO) on the same workload."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    import pic_configs as cfg

    t0 = time.perf_counter()
    sim = O.Sim(w.NX, w.NY, w.width, w.height, w.dt, cfg.EPS, w.bz, w.rects(), bx=w.bx, deposit_energy=1,
                use_pardiso=True)
    for s, (q, m, _) in enumerate(w.species()):
        sim.add_species(q, m, w.n_per_species, cfg.N0, *parts[s])
    sim.step()                      # warm-up (first solve after factorisation)
    sim.times()
    n0 = sum(sim.count(k) for k in range(2))
    t1 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    t2 = time.perf_counter()
    n1 = sum(sim.count(k) for k in range(2))
    stages = sim.times()
    pardiso = sim.uses_pardiso
    sim.close()
    rate = steps * 0.5 * (n0 + n1) / (t2 - t1)
    log(f"cpu baseline: {rate:.3e} particle-steps/s over {steps} steps ({t2 - t1:.2f} s, setup {t1 - t0:.1f} s)")
    return {
        "value": rate, "unit": "particle-steps/s", "cores": O.threads(), "kind": "port",
        "sample": f"full {w.name} workload ({w.NX}x{w.NY} cells, {2 * w.n_per_species} particles), {steps} steps "
                  f"after 1 warm-up step; FP64 AoS, OpenMP, Poisson by "
                  f"{'MKL PARDISO (factorised before timing)' if pardiso else 'the oracle direct solve'}",
        "stage_s": {k: round(v / steps, 5) for k, v in stages.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="electrode_1024")
    ap.add_argument("--layout", type=int, default=None, help="0 F64, 1 CELL32 (default: the workload's)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--learned-train", type=int, default=2000,
                    help="online training steps of the learned-guess comparison (0 skips it)")
    ap.add_argument("--learned-steps", type=int, default=5000,
                    help="steps of the learned-guess comparison, training included")
    ap.add_argument("--learned-grids", default="256,1024",
                    help="grids (cells per side, electrode geometry) of the learned-guess comparison")
    ap.add_argument("--precision-steps", type=int, default=20,
                    help="steps of the precision-form leg (0 skips it)")
    ap.add_argument("--ensemble-runs", type=int, default=64,
                    help="runs of the batched-ensemble comparison (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--published-steps", type=int, default=20,
                    help="steps of the reference's published periodic 501 x 257 run (0: skip)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="Poisson: FP64 PCG, or FP32 correction solves refined in FP64")
    ap.add_argument("--guess", choices=("prev", "extrap"), default="prev",
                    help="Poisson initial guess: previous phi, or extrapolated from the last two")
    ap.add_argument("--smoother", choices=("chebyshev", "jacobi"), default="chebyshev",
                    help="multigrid smoother of the Poisson preconditioner")
    ap.add_argument("--injection-steps", type=int, default=10,
                    help="steps timed with particle injection on (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three MIB-MiB device copies (counter calibration) and exit")
    ap.add_argument("--unet-only", type=int, default=0, metavar="N",
                    help="only time the learned-guess network on an N x N image (forward, forward + backward) "
                         "and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.unet_only:
        unet_only(a.unet_only)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1", "n_gpus": a.gpus}))
        sys.exit(2)

    import numpy as np

    import pic_configs as cfg
    import pic_hip

    w = cfg.WORKLOADS[a.workload]
    layout = w.layout if a.layout is None else a.layout
    t0 = time.perf_counter()
    parts = [cfg.generate(w, s) for s in range(2)]
    log(f"generated {2 * w.n_per_species} particles in {time.perf_counter() - t0:.1f} s")
    ctx = cfg.make_context(w, layout=layout, particles=parts, poisson_tol=1e-10,
                           poisson_guess=1 if a.guess == "extrap" else 0,
                           poisson_precision=1 if a.precision == "mixed" else 0,
                           mg_smoother=1 if a.smoother == "chebyshev" else 0, capacity_factor=1.05)
    log(f"context ready ({time.perf_counter() - t0:.1f} s)")

    it = 0
    for _ in range(a.warmup):
        ctx.step(it)
        it += 1
    ctx.sync()
    n_start = sum(ctx.count(s) for s in range(2))
    # timed loop: the scatter kernel (the dominant one) carries dispatch
    # timestamps, no stream markers; the other particle kernels and the
    # per-stage breakdown are measured over separate steps after it
    ctx.set_timing(3)
    ctx.stage_times()
    stats = []
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        stats.append(ctx.step(it))
        it += 1
    ctx.sync()
    t2 = time.perf_counter()
    live_times = ctx.stage_times()
    n_end = sum(ctx.count(s) for s in range(2))
    nb = min(10, a.steps)
    ctx.set_timing(2)
    ctx.stage_times()
    for _ in range(nb):
        ctx.step(it)
        it += 1
    kernel_times = ctx.stage_times()
    ctx.set_timing(1)
    for _ in range(nb):
        ctx.step(it)
        it += 1
    stage_ms = {k: v[0] / nb for k, v in ctx.stage_times().items()}
    ctx.set_timing(0)
    elapsed = t2 - t1
    ps = a.steps * 0.5 * (n_start + n_end)
    value = ps / elapsed
    ms_step = 1e3 * elapsed / a.steps
    log(f"gpu: {value:.3e} particle-steps/s, {ms_step:.3f} ms/step")

    # roofline of the dominant particle kernel (one launch per step serves
    # both species), its launch durations taken over the timed loop itself
    stages = kernel_times
    dom = max(KERNEL_BYTES, key=lambda k: stages[k][0])
    live = bool(live_times[dom][1])
    k_ms, k_n = live_times[dom] if live else stages[dom]
    avg_s = k_ms / max(k_n, 1) * 1e-3
    n_avg = 0.5 * (n_start + n_end)                # particles of both species per step
    nodes = (w.NX + 1) * (w.NY + 1)

    def launch_bytes(kind, launches, steps):
        # algorithmic bytes of one launch: the step's particles and both
        # species' cells, over the launches the step makes
        pp, pc = KERNEL_BYTES[kind][layout]
        return (pp * n_avg + pc * 2 * nodes) * steps / max(launches, 1)

    per_p, per_c = KERNEL_BYTES[dom][layout]
    kbytes = launch_bytes(dom, k_n, a.steps if live else nb)
    achieved = kbytes / avg_s / 1e9 if avg_s > 0 else None
    traffic = None
    tf = ROOT / "profiles" / f"{KERNEL_NAMES[dom]}_traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    roof = {"kernel": KERNEL_NAMES[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
            "bytes_per_launch": kbytes, "bytes_per_particle": per_p, "bytes_per_cell": per_c,
            "avg_launch_us": avg_s * 1e6, "launches_timed": k_n,
            "launches_per_step": k_n / (a.steps if live else nb),
            "timed_over": "the timed loop" if live else "separate steps after the timed loop"}
    others = {}
    for k in KERNEL_BYTES:
        ms_k, n_k = stages[k]
        if n_k:
            b = launch_bytes(k, n_k, nb)
            others[KERNEL_NAMES[k]] = {"avg_launch_us": round(ms_k / n_k * 1e3, 2),
                                       "GB_s": round(b / (ms_k / n_k * 1e-3) / 1e9, 1)}

    injection = None
    if a.injection_steps > 0:
        try:
            injection, it = injection_form(ctx, w, it, a.injection_steps)
            log(f"with injection: {injection['ms_per_step']:.3f} ms/step")
        except Exception as e:  # reported, the main measurement stands
            log(f"injection run failed: {e}")
            injection = {"error": str(e)}

    learned = None
    if a.learned_train > 0:
        learned = {}
        base = cfg.WORKLOADS["electrode_1024"]
        for g in (int(v) for v in a.learned_grids.split(",") if v):
            try:
                lw = base if g == base.NX else base.scaled(NX=g, n=base.n_per_species * g * g // (base.NX * base.NX),
                                                           name=f"electrode_{g}")
                lparts = parts if lw is w else [cfg.generate(lw, s) for s in range(2)]
                learned[str(g)] = learned_vs_plain(lw, lparts, 1, a.learned_train, a.learned_steps)
            except Exception as e:  # reported, the main measurement stands
                log(f"learned-guess comparison on {g}^2 failed: {e}")
                learned[str(g)] = {"error": str(e)}

    precision = None
    if a.precision_steps > 0:
        try:
            precision = precision_form(a.precision_steps)
        except Exception as e:  # reported, the main measurement stands
            log(f"precision-form leg failed: {e}")
            precision = {"error": str(e)}

    ensemble = None
    if a.ensemble_runs > 0:
        try:
            ensemble = ensemble_form(a.ensemble_runs, a.ensemble_steps)
        except Exception as e:  # reported, the main measurement stands
            log(f"ensemble comparison failed: {e}")
            ensemble = {"error": str(e)}

    published = None
    if a.published_steps > 0:
        try:
            published = published_workload(a.published_steps)
        except Exception as e:  # reported, the main measurement stands
            log(f"published-workload leg failed: {e}")
            published = {"error": str(e)}

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(w, parts, a.cpu_steps)
        except Exception as e:  # the GPU result stands without the baseline
            log(f"cpu baseline failed: {e}")
    iters = [s["iterations"] for s in stats]
    out = {
        "metric": "particle-steps/s (16M ptcl, 1024\u00b2); Poisson ms learned vs plain @1e-8 res",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "int32 cell + f32 offsets/velocities" if layout == 1 else "f64",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": w.name, "cells": f"{w.NX}x{w.NY}", "particles_per_species": w.n_per_species,
                   "species": 2, "electrode_V": w.electrode_v if w.electrode else None,
                   "B_T": w.b_total, "B_tilt_deg": w.b_tilt_deg, "poisson_tol": 1e-10,
                   "poisson_guess": a.guess, "poisson_precision": a.precision, "mg_smoother": a.smoother},
        "stages_ms_per_step": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_breakdown_note": "per-stage HIP event markers, measured over separate steps after the timed loop",
        "poisson": {"iterations_mean": float(np.mean(iters)), "ms_mean": float(np.mean([s["ms"] for s in stats])),
                    "inner_iterations_mean": float(np.mean([s["inner_iterations"] for s in stats])),
                    "initial_relres_mean": float(np.mean([s["initial_relres"] for s in stats]))},
        "roofline": roof,
        "particle_kernels": others,
        "with_injection": injection,
        "poisson_learned_vs_plain": learned,
        "precision_form": precision,
        "ensemble": ensemble,
        "reference_published_workload": published,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
