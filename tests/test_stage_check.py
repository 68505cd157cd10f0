"""The set comparison the stage checks use: rows permuted inside their cells
compare equal; one flipped bit in any field, a row moved to another cell or a
missing row do not."""
import numpy as np

import stage_check as SC


def _rows(n=20000, seed=0):
    rng = np.random.default_rng(seed)
    cell = np.sort(rng.integers(0, 500, n)).astype(np.int32)
    f32 = [rng.random(n, dtype=np.float32) for _ in range(5)]
    return (cell, *f32)


def test_permutation_inside_cells_is_equal():
    a = _rows()
    perm = np.lexsort((np.random.default_rng(1).random(a[0].size), a[0]))
    b = tuple(x[perm] for x in a)
    assert SC.same_particles(a, b)
    f64 = (a[0], *[x.astype(np.float64) for x in a[1:]])
    assert SC.same_particles(f64, tuple(x[perm] for x in f64))


def test_any_single_change_is_caught():
    a = _rows()
    for field in range(1, 6):
        b = [x.copy() for x in a]
        u = b[field].view(np.uint32)
        u[1234] ^= 1
        assert not SC.same_particles(a, tuple(b)), field
    b = [x.copy() for x in a]
    b[0][77] += 1
    assert not SC.same_particles(a, tuple(b))
    assert not SC.same_particles(a, tuple(x[:-1] for x in a))
    # two rows exchanging one field (same multiset per column, different rows)
    b = [x.copy() for x in a]
    i, j = 10, 11
    b[3][i], b[3][j] = a[3][j], a[3][i]
    assert not SC.same_particles(a, tuple(b))
