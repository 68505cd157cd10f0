"""The FP64 matrix-core GEMM of the direct Poisson solve (k_dgemm) against
NumPy: every operand orientation, the three epilogues (store, elementwise
scale, accumulate) and shapes that are not multiples of the 32 x 32 tile or
the 32-deep k chunk, including the transform shapes of the 501 x 257-node
periodic grid (499 interior columns, period 256)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 4), (1, 1, 1), (37, 53, 29), (64, 33, 100), (256, 499, 499), (499, 256, 256)]


@pytest.mark.parametrize("M,N,K", SHAPES)
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_dgemm_orientations(hip_mod, M, N, K, ta, tb):
    rng = np.random.default_rng(M * 1000 + N + K)
    a = rng.standard_normal((K, M) if ta else (M, K))
    b = rng.standard_normal((N, K) if tb else (K, N))
    ref = (a.T if ta else a) @ (b.T if tb else b)
    got = hip_mod.dgemm_check(a, b, ta=ta, tb=tb)
    # FP64 with a different summation order: |error| <= K eps sum |a||b|
    bound = 4 * K * np.finfo(np.float64).eps * (np.abs(a.T if ta else a) @ np.abs(b.T if tb else b))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


@pytest.mark.parametrize("M,N,K", [(40, 70, 90), (256, 499, 256)])
def test_dgemm_epilogues(hip_mod, M, N, K):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((M, K))
    b = rng.standard_normal((K, N))
    s = rng.uniform(0.5, 2.0, (M, N))
    c0 = rng.standard_normal((M, N))
    ref = a @ b
    bound = 4 * K * np.finfo(np.float64).eps * (np.abs(a) @ np.abs(b))
    got = hip_mod.dgemm_check(a, b, scale=s)
    assert np.all(np.abs(got - ref * s) <= bound * s + 1e-300)
    got = hip_mod.dgemm_check(a, b, c=c0)
    assert np.all(np.abs(got - (c0 + ref)) <= bound + 4 * np.finfo(np.float64).eps * np.abs(c0 + ref))
    # the same call twice gives the same bits (fixed summation order)
    assert np.array_equal(hip_mod.dgemm_check(a, b), hip_mod.dgemm_check(a, b))
