"""pcg_count_cells_noalloc (the counting pass with its scratch in a fixed
device array) against pcg_count_cells, through the C ABI: two calls in a row,
the second on fewer points with a smaller grid, then an out-of-box input and
a good call after it (the error flag must not stick)."""
import ctypes

import numpy as np
import pytest

import potreeconverter_amd as pc

pytestmark = pytest.mark.gpu


def _dp(v):
    return (ctypes.c_double * 3)(*v)


def _grids(L, d, n, recs, G, shrink=1.0):
    rl = 26
    a = np.frombuffer(recs, dtype=np.uint8)[: n * rl].reshape(n, rl)[:, :12].copy().view(np.int32).reshape(n, 3)
    scale, off = [0.001] * 3, [0.0] * 3
    mn = [a[:, k].min() * 0.001 for k in range(3)]
    mx = [mn[k] + (a[:, k].max() * 0.001 - mn[k]) * shrink for k in range(3)]
    out = []
    for fn in (L.pcg_count_cells, L.pcg_count_cells_noalloc):
        dg = ctypes.c_void_p()
        assert L.pcg_device_alloc(ctypes.byref(dg), 4 * G ** 3) == 0
        # stale contents: the entry point must clear the grid itself
        junk = np.full(G ** 3, 7, dtype=np.int32)
        assert L.pcg_copy(dg, junk.ctypes.data_as(ctypes.c_void_p), 4 * G ** 3, 0) == 0
        rc = fn(d, n, rl, _dp(scale), _dp(off), _dp(scale), _dp(mn), _dp(mn), _dp(mx), G, dg, None)
        got = np.zeros(G ** 3, dtype=np.int32)
        if rc == 0:
            assert L.pcg_copy(got.ctypes.data_as(ctypes.c_void_p), dg, 4 * G ** 3, 1) == 0
        assert L.pcg_device_free(dg) == 0
        out.append((rc, got))
    return out


def test_noalloc_count_matches_allocating_entry_point():
    L = pc.lib()
    assert L.pcg_init() == 0
    n = 400_000
    recs = pc.LasGen(0, n, seed=12).records()
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), len(recs) + 64) == 0
    assert L.pcg_copy(d, recs, len(recs), 0) == 0
    try:
        for k, (m, G) in enumerate(((n, 128), (150_000, 64))):
            print(f"noalloc call {k}: n={m} grid={G}", flush=True)
            (rc0, a), (rc1, b) = _grids(L, d, m, recs, G)
            assert rc0 == 0 and rc1 == 0, pc.lib().pcg_last_error()
            assert a.sum() == m and np.array_equal(a, b)
        print("noalloc out-of-box call", flush=True)
        (rc0, _), (rc1, _) = _grids(L, d, 100_000, recs, 32, shrink=0.5)
        assert rc0 == 123 and rc1 == 123
        print("noalloc call after the error", flush=True)
        (rc0, a), (rc1, b) = _grids(L, d, 200_000, recs, 128)
        assert rc0 == 0 and rc1 == 0 and np.array_equal(a, b) and b.sum() == 200_000
    finally:
        L.pcg_device_free(d)


def test_noalloc_rejects_sizes_it_cannot_take():
    L = pc.lib()
    d = ctypes.c_void_p()
    z = _dp([0.0] * 3)
    one = _dp([1.0] * 3)
    assert L.pcg_count_cells_noalloc(d, 10, 26, one, z, one, z, z, one, 1024, d, None) != 0
    assert b"grid_size" in L.pcg_last_error()
    assert L.pcg_count_cells_noalloc(d, 10, 200, one, z, one, z, z, one, 128, d, None) != 0
