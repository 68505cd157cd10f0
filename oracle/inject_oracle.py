"""NumPy restatement of particle injection (bmp-6th-sem_amd/csrc/pic_inject.hip).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU injection.

What the reference does (SYCL-PIC-poissonsolver/src):
  * main.cpp:235-238 — after both movers, num_new = (electrons that left
    through x < 0) - (ions that did), clamped at 0 (the exits are counted in
    mover.cpp:344-352);
  * utils.cpp:374-499 generate_new_electrons — num_new electrons at x = 0.024,
    y = GRID_HEIGHT * U, velocities v * sqrt(-ln U) * (cos, sin)(2 pi U) with
    v = sqrt(2 |q/m| electronTemp), electronTemp passed as an int (:381);
  * utils.cpp:610-703 generate_pairs — fn = GRID_HEIGHT * dt * (2/PI) *
    (x22 - x11) * ionimax / 1.67e6 with PI = 3.141592654 (:17), ionimax =
    5.23e23 (:619); target = ceil(fn); up to max(10 target, 100) attempts of
    x = xm + asin(2U - 1) (x22 - x11) / PI, y = GRID_HEIGHT * U, kept when
    0 < x < GRID_WIDTH, until target pairs; electron at electronTemp, ion at
    ionTemp, same position.
The reference draws from time-seeded std::rand / ran2.  This build draws from
a counter-based generator instead (same distributions, reproducible):
  u(seed, step, stream, k, j) = ((z >> 12) + 1/2) 2^-52,
  z = mix(mix(seed + G (step + 1)) ^ (stream << 60) ^ (k << 8) ^ j),
with mix the SplitMix64 output function and G = 0x9E3779B97F4A7C15.
Normals come from the Box-Muller transform r = sqrt(-2 ln u), (r cos 2 pi v,
r sin 2 pi v); a Maxwellian uses draws j..j+3: (vx, vy) from the first pair,
vz the cosine of the second, each times sigma = sqrt(|q/m| T).
Stream 0: re-injected electron k: y from draw 0, velocity from draws 1-4.
Stream 1: pair attempt a: x from draw 0, y from draw 1, electron velocity
draws 2-5, ion velocity draws 6-9.
"""
from __future__ import annotations

import math

import numpy as np

PI_REF = 3.141592654            # utils.cpp:17
GOLDEN = 0x9E3779B97F4A7C15
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    """SplitMix64 output function on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, n):
    """First n outputs of the SplitMix64 generator started at `seed`."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.uint64(GOLDEN) * np.arange(1, n + 1, dtype=np.uint64)
    return mix64(s)


def urand(seed, step, stream, k, j):
    """Uniform in (0, 1) for indices k (array) and draw j."""
    with np.errstate(over="ignore"):
        base = mix64(np.uint64(seed) + np.uint64(GOLDEN) * np.uint64(step + 1))
        z = mix64(base ^ (np.uint64(stream) << np.uint64(60)) ^ (np.asarray(k, np.uint64) << np.uint64(8))
                  ^ np.uint64(j))
    return ((z >> np.uint64(12)).astype(np.float64) + 0.5) * 2.0 ** -52


def normal2(seed, step, stream, k, j):
    u = urand(seed, step, stream, k, j)
    v = urand(seed, step, stream, k, j + 1)
    r = np.sqrt(-2.0 * np.log(u))
    t = 6.283185307179586 * v
    return r * np.cos(t), r * np.sin(t)


def maxwell(seed, step, stream, k, j, sigma):
    a, b = normal2(seed, step, stream, k, j)
    c, _ = normal2(seed, step, stream, k, j + 2)
    return sigma * a, sigma * b, sigma * c


def pair_rate(height, dt, x11=0.0025, x22=0.01, ionimax=5.23e23):
    """fn of generate_pairs (utils.cpp:619-620), in the reference's order."""
    return height * dt * (2.0 / PI_REF) * (x22 - x11) * ionimax * (1.0 / (1.67e6))


def pair_target(height, dt, x11=0.0025, x22=0.01, ionimax=5.23e23):
    """(target, attempts) of generate_pairs (utils.cpp:621-635)."""
    fn = pair_rate(height, dt, x11, x22, ionimax)
    if fn <= 0.0:
        return 0, 0
    t = int(math.ceil(fn))
    return t, max(10 * t, 100)


def reinject_count(left_e, left_i):
    """num_new of main.cpp:235-238."""
    return max(0, int(left_e) - int(left_i))


def _y(height, u):
    y = height * u
    return np.where(y < height, y, 0.0)


def reinjected(seed, step, n, x_inject, height, sigma):
    """The n re-injected electrons of injection `step`: (x, y, vx, vy, vz)."""
    k = np.arange(n, dtype=np.uint64)
    y = _y(height, urand(seed, step, 0, k, 0))
    vx, vy, vz = maxwell(seed, step, 0, k, 1, sigma)
    return np.full(n, float(x_inject)), y, vx, vy, vz


def pair_x(u, x11=0.0025, x22=0.01, xm=0.00625):
    return xm + np.arcsin(2.0 * u - 1.0) * (x22 - x11) / PI_REF


def pairs(seed, step, target, attempts, width, height, sig_e, sig_i, x11=0.0025, x22=0.01, xm=0.00625):
    """The pairs of injection `step`: the first `target` attempts (in order)
    with 0 < x < width.  Returns (electrons, ions), each (x, y, vx, vy, vz)."""
    a = np.arange(attempts, dtype=np.uint64)
    x = pair_x(urand(seed, step, 1, a, 0), x11, x22, xm)
    keep = a[(x > 0.0) & (x < width)][:target]
    x = x[keep.astype(np.int64)]
    y = _y(height, urand(seed, step, 1, keep, 1))
    e = (x, y, *maxwell(seed, step, 1, keep, 2, sig_e))
    i = (x.copy(), y.copy(), *maxwell(seed, step, 1, keep, 6, sig_i))
    return e, i


def injection(seed, step, left_e, left_i, *, width, height, dt, qm_e, qm_i, x_inject=0.024, te_reinject=10.0,
              te=10.0, ti=0.026, x11=0.0025, x22=0.01, xm=0.00625, ionimax=5.23e23, reinject=True,
              pairs_on=True, max_reinject=1 << 20):
    """All particles of injection `step` before any capacity clamp:
    (electrons, ions) as (x, y, vx, vy, vz), electrons = re-injected then pair
    electrons (the staging order)."""
    n = reinject_count(left_e, left_i) if reinject else 0
    if not (0.0 <= x_inject < width):
        n = 0
    n = min(n, max_reinject)
    re = reinjected(seed, step, n, x_inject, height, math.sqrt(abs(qm_e) * te_reinject))
    if pairs_on:
        target, attempts = pair_target(height, dt, x11, x22, ionimax)
    else:
        target, attempts = 0, 0
    pe, pi = pairs(seed, step, target, attempts, width, height, math.sqrt(abs(qm_e) * te),
                   math.sqrt(abs(qm_i) * ti), x11, x22, xm)
    e = tuple(np.concatenate([a, b]) for a, b in zip(re, pe))
    return e, pi
