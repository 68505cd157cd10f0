"""The learned-guess network on the GPU (hand-written BF16 MFMA kernels) against
the NumPy oracle (oracle/unet_oracle.py): MFMA lane layout, forward pass,
gradients, Adam step, training, and its use as the Poisson initial guess.

Tolerances: the GPU sums in FP32 in another order than the oracle's FP64 and
both round to BF16 at the same points, so an activation can differ by one BF16
unit where a sum lands near a rounding tie; outputs are compared normwise to
2e-2, gradients per parameter block normwise to 5e-2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 128


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _inputs(seed=0, n=N):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n] / n
    x = np.stack([np.sin(6 * xx) * np.cos(4 * yy) + 0.3 * rng.normal(size=(n, n)),
                  np.cos(3 * xx + yy) + 0.1 * rng.normal(size=(n, n))], axis=-1)
    return x.astype(np.float32)


def _params(net, seed=0):
    import unet_oracle as U
    p = net.params().astype(np.float64)
    _, hw, hb, _ = U.layout()
    rng = np.random.default_rng(seed)
    p[hw:hw + 16] = rng.normal(0, 0.3, 16)        # a head large enough to pass gradients
    p[hb] = 0.1
    net.set_params(p.astype(np.float32))
    return net.params().astype(np.float64)


def test_mfma_lane_layout(hip_mod):
    rng = np.random.default_rng(1)
    a = rng.integers(-8, 8, (16, 16)).astype(np.float32)
    b = rng.integers(-8, 8, (16, 16)).astype(np.float32)
    d = hip_mod.mfma_probe(a, b)
    assert np.array_equal(d, a @ b)


def test_mfma32_lane_layout(hip_mod):
    """The forward convolutions' 16x16x32 instruction: lane l holds A[l%16][8(l/16)+e]
    and B[8(l/16)+e][l%16] and receives D[4(l/16)+e][l%16]."""
    rng = np.random.default_rng(2)
    a = rng.integers(-8, 8, (16, 32)).astype(np.float32)
    b = rng.integers(-8, 8, (32, 16)).astype(np.float32)
    d = hip_mod.mfma32_probe(a, b)
    assert np.array_equal(d, a @ b)


@pytest.mark.parametrize("n", [N, 1024])
def test_unet_forward_matches_oracle(hip_mod, n):
    """Inference forward pass (fused pooling, upsampling read in place, fused
    head) at 128 (images narrower than a 64-pixel tile on the coarse levels)
    and 1024 (the learned guess's image, two channel tiles per workgroup on
    the 512 level), and the training forward pass (every activation kept),
    both against the oracle."""
    import unet_oracle as U
    net = hip_mod.UNet(n, seed=7)
    p = _params(net, 1)
    x = _inputs(1, n)
    out = net.forward(x)
    ref, _ = U.forward(p, x.astype(np.float64))
    print(f"forward error at {n}: {_rel(out, ref):.3e}")
    assert _rel(out, ref) < 2e-2
    if n == N:
        tgt = np.zeros((n, n), np.float32)
        loss = net.train_step(x, tgt, np.ones((n, n), np.uint8), 0.0)
        assert loss == pytest.approx(float(np.mean(ref ** 2)), rel=2e-2)
    net.close()


def test_unet_gradients_match_oracle(hip_mod):
    import unet_oracle as U
    net = hip_mod.UNet(N, seed=8)
    p = _params(net, 2)
    x = _inputs(2)
    rng = np.random.default_rng(3)
    tgt = (0.5 * x[:, :, 0] + 0.2 * rng.normal(size=(N, N))).astype(np.float32)
    mask = np.ones((N, N), np.uint8)
    mask[0, :] = mask[:, 0] = 0
    loss = net.train_step(x, tgt, mask, 0.0)
    g = net.grads().astype(np.float64)
    rloss, rg = U.loss_and_grads(p, x.astype(np.float64), tgt.astype(np.float64), mask.astype(np.float64))
    assert loss == pytest.approx(rloss, rel=1e-2)
    blocks, hw, hb, n = U.layout()
    errs = []
    for k, (w, b, ci, co) in enumerate(blocks):
        errs.append((k, _rel(g[w:w + 9 * co * ci], rg[w:w + 9 * co * ci]), _rel(g[b:b + co], rg[b:b + co])))
    errs.append(("head", _rel(g[hw:hw + 17], rg[hw:hw + 17]), 0.0))
    print("gradient errors (block, W, b):", errs)
    assert all(e[1] < 5e-2 and e[2] < 5e-2 for e in errs), errs
    net.close()


def test_unet_adam_step_and_training(hip_mod):
    import unet_oracle as U
    net = hip_mod.UNet(N, seed=9)
    p0 = _params(net, 3).astype(np.float32)
    x = _inputs(4)
    tgt = (0.25 * (np.roll(x[:, :, 0], 1, 0) + np.roll(x[:, :, 0], -1, 1)) - 0.3 * x[:, :, 1]).astype(np.float32)
    mask = np.ones((N, N), np.uint8)
    # one step against the oracle's Adam on the gradients that step used (the
    # weight-gradient sums use FP32 atomics, so a recomputation may differ in
    # the last bits, and Adam's first step divides g by |g|)
    net.train_step(x, tgt, mask, 1e-3)
    p1 = net.params()
    g = net.grads()
    pr, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    U.adam(pr, g, m, v, 1, 1e-3)
    assert np.abs(p1 - pr).max() <= 1e-6 * max(1.0, np.abs(pr).max())
    # training lowers the loss
    losses = [net.train_step(x, tgt, mask, 1e-3) for _ in range(60)]
    print("loss first/last:", losses[0], losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    net.close()


def test_learned_guess_in_the_solver(hip_mod):
    """256 x 256 cells, electrode geometry, tolerance 1e-8: the PCG with the
    learned guess reaches the same potential as the plain one.  After 40
    training steps the trained network earns its place: its guess is kept on
    at least 80 % of the solves it is tried on, and over the last 20 solves it
    starts the PCG from at most half the plain warm start's relative residual
    (measured on MI355X: 0.0142 against 0.0396, used 39 of 39), with no more
    iterations on average."""
    import pic_configs as cfg
    w = cfg.WORKLOADS["electrode_512"].scaled(NX=256, n=200_000, name="learned_256")
    parts = [cfg.generate(w, s) for s in range(2)]
    # both on the PCG (the direct solve would leave the plain run nothing to iterate)
    a = cfg.make_context(w, particles=parts, poisson_tol=1e-8, poisson_direct=0)
    b = cfg.make_context(w, particles=parts, poisson_tol=1e-8, poisson_direct=0)
    b.learned_enable(train_steps=40, min_trained=20, lr=1e-3, seed=1)
    sa, sb = [], []
    for it in range(60):
        sa.append(a.step(it))
        sb.append(b.step(it))
        pa, pb = a.grid(hip_mod.GRID_PHI), b.grid(hip_mod.GRID_PHI)
        assert np.abs(pa - pb).max() <= 1e-6 * np.abs(pa).max(), it
    st = b.learned_stats()
    print("learned stats:", st)
    tail = slice(40, 60)
    ia = np.mean([s["iterations"] for s in sa[tail]])
    ib = np.mean([s["iterations"] for s in sb[tail]])
    ra = np.mean([s["initial_relres"] for s in sa[tail]])
    rb = np.mean([s["initial_relres"] for s in sb[tail]])
    print(f"iterations plain {ia:.2f} learned {ib:.2f}; initial relres plain {ra:.3e} learned {rb:.3e}")
    assert st["trained"] == 40 and st["tried"] == 60 - 1 - 20 and st["used"] >= 0.8 * st["tried"]
    assert all(s["final_relres"] <= 1e-8 for s in sb)
    assert rb <= 0.5 * ra, (ra, rb)
    assert ib <= ia, (ia, ib)
    a.close()
    b.close()


def test_learned_guess_replaces_the_extrapolated_one(hip_mod):
    """With poisson_guess = 1 (linear extrapolation) and the learned guess on,
    the learned guess is the one offered once the network has trained (the
    extrapolation no longer overrides it): the guess is tried and used, and
    every solve reaches the same potential as a plain run at 1e-8."""
    import pic_configs as cfg
    w = cfg.WORKLOADS["electrode_512"].scaled(NX=256, n=100_000, name="learned_extrap_256")
    parts = [cfg.generate(w, s) for s in range(2)]
    a = cfg.make_context(w, particles=parts, poisson_tol=1e-8, poisson_direct=0)
    b = cfg.make_context(w, particles=parts, poisson_tol=1e-8, poisson_direct=0, poisson_guess=1)
    b.learned_enable(train_steps=12, min_trained=8, lr=1e-3, seed=2)
    for it in range(20):
        a.step(it)
        sb = b.step(it)
        assert sb["final_relres"] <= 1e-8, (it, sb)
        pa, pb = a.grid(hip_mod.GRID_PHI), b.grid(hip_mod.GRID_PHI)
        assert np.abs(pa - pb).max() <= 1e-6 * np.abs(pa).max(), it
    st = b.learned_stats()
    print("learned stats with extrapolation requested:", st)
    assert st["tried"] == 20 - 1 - 8 and st["used"] >= 1
    a.close()
    b.close()


def test_unet_residual_loss_gradients_match_oracle(hip_mod):
    import unet_oracle as U
    net = hip_mod.UNet(N, seed=10)
    p = _params(net, 5)
    x = _inputs(5)
    rng = np.random.default_rng(6)
    r = rng.uniform(-1, 1, (N, N))
    mask = np.ones((N, N), np.uint8)
    mask[0, :] = mask[:, 0] = 0
    loss = net.train_step_residual(x, r, mask, 0.25, 0.25, 0.0)
    g = net.grads().astype(np.float64)
    rloss, rg = U.residual_loss_and_grads(p, x.astype(np.float64), r, mask.astype(np.float64), 0.25, 0.25)
    assert loss == pytest.approx(rloss, rel=1e-2)
    assert _rel(g, rg) < 5e-2
    net.close()


def test_training_is_reproducible(hip_mod):
    """Two networks from the same seed trained on the same data end every step
    with the same loss, gradients and parameters, bit for bit: the weight,
    bias, head and loss reductions add per-wave / per-workgroup partials in a
    fixed order (no floating-point atomics)."""
    nets = [hip_mod.UNet(N, seed=3) for _ in range(2)]
    for net in nets:
        _params(net, 4)
    x = _inputs(4)
    rng = np.random.default_rng(9)
    tgt = rng.normal(0, 0.2, (N, N)).astype(np.float32)
    r = rng.uniform(-1, 1, (N, N))
    mask = np.ones((N, N), np.uint8)
    mask[0, :] = mask[:, 0] = 0
    for k in range(3):
        la = nets[0].train_step(x, tgt, mask, 1e-3)
        lb = nets[1].train_step(x, tgt, mask, 1e-3)
        assert la == lb, k
        assert np.array_equal(nets[0].grads(), nets[1].grads()), k
        assert np.array_equal(nets[0].params(), nets[1].params()), k
        la = nets[0].train_step_residual(x, r, mask, 0.25, 0.25, 1e-3)
        lb = nets[1].train_step_residual(x, r, mask, 0.25, 0.25, 1e-3)
        assert la == lb, k
        assert np.array_equal(nets[0].grads(), nets[1].grads()), k
    # the same network on the same data twice (no update): the same gradients
    g0 = (nets[0].train_step(x, tgt, mask, 0.0), nets[0].grads())
    g1 = (nets[0].train_step(x, tgt, mask, 0.0), nets[0].grads())
    assert g0[0] == g1[0] and np.array_equal(g0[1], g1[1])
    for net in nets:
        net.close()
