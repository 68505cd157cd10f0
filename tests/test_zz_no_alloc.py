"""pic_bin_keys_noalloc, the counting sort that allocates nothing, against the
bin stage (pic_bin, through pic_upload_particles) on the same keys: counts and
starts equal bit for bit, slots a permutation that puts every key inside its
cell's range, dead keys -1.  Two calls in a row share the library's fixed scan
scratch, the second on a smaller input (a 2048 x 1101-node grid, whose chunk
totals are scanned on their own, then a 129 x 129 one).  The test's own device
buffers come from the HIP runtime directly."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, ME = 1.602176634e-19, 9.1093837015e-31
WALLS = [[0, 0, 0, 1, 0.0], [1, 0, 1, 1, 0.0], [0, 1, 1, 1, 0.0], [0, 0, 1, 0, 0.0]]


class _Hip:
    """Device buffers for the test (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(max(arr.nbytes, 16))), "hipMalloc")
        self.bufs.append(p)
        if arr.nbytes:
            self._ok(self.rt.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1), "hipMemcpy")
        return p

    def empty(self, n, dtype=np.int32):
        return self.put(np.full(n, -7, dtype))

    def get(self, p, n, dtype=np.int32):
        out = np.empty(n, dtype)
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        if n:
            self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _bind(lib):
    f = lib.pic_bin_keys_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def _case(hip, dev, f, NX, NY, n, seed):
    dx = 1e-4
    p = hip.default_params(NX=NX, NY=NY, width=NX * dx, height=NY * dx, dt=1.77e-11, layout=1)
    c = hip.Context(p)
    c.set_geometry(WALLS)
    s = c.add_species(-E, ME, n, 1e16, n)
    x, y, vx, vy, vz = hip.generate_particles(seed, 0, n, c.params.width, c.params.height, c.dx, c.dy, E / ME, 2.0)
    c.upload(s, x, y, vx, vy, vz)                  # bins the species (pic_bin's counting sort)
    counts_ref, start_ref = c.cell_counts(s), c.cell_start(s)
    cell = c.download_cell32(s)[0]
    G = c.G
    c.close()
    rng = np.random.default_rng(seed)
    dead = np.array([-1] * 37 + [G, G + 5, -100], np.int32)
    keys = rng.permutation(np.concatenate([cell, dead])).astype(np.int32)
    dk = dev.put(keys)
    dc, ds, dsl = dev.empty(G), dev.empty(G + 1), dev.empty(keys.size)
    assert f(dk, keys.size, G, dc, ds, dsl, None) == 0, hip.load().pic_last_error()
    counts, start, slot = dev.get(dc, G), dev.get(ds, G + 1), dev.get(dsl, keys.size)
    assert np.array_equal(counts, counts_ref)
    assert np.array_equal(start, start_ref)
    live = (keys >= 0) & (keys < G)
    assert np.all(slot[~live] == -1)
    assert np.array_equal(np.sort(slot[live]), np.arange(live.sum()))
    k = keys[live]
    assert np.all((slot[live] >= start_ref[k]) & (slot[live] < start_ref[k + 1]))
    return G


def test_bin_keys_noalloc_matches_pic_bin(hip_mod):
    hip = hip_mod
    f = _bind(hip.load())
    dev = _Hip()
    try:
        G1 = _case(hip, dev, f, 2047, 1100, 400000, 3)
        G2 = _case(hip, dev, f, 128, 128, 20000, 4)
        assert G2 < G1
        # sizes beyond the fixed scratch are refused before any work
        z = dev.empty(16)
        assert f(z, 0, (1 << 24) + 1, z, z, z, None) == -11
        assert f(z, -1, 16, z, z, z, None) == -11
        assert f(z, 4, 16, None, z, z, None) == -1
    finally:
        dev.free()
