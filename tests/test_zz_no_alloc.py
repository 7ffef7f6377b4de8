"""The allocation-free scaled-overlap scan (hsmc_gpu_any_overlap_scaled_noalloc)
against the allocating entry point it mirrors (hsmc_gpu_any_overlap_scaled),
both called through the C ABI: identical flags on two calls in a row, the
second on a smaller replica with fewer scale factors (the shared device
array still holds the first call's larger grid), then the first call again;
and a replica too large for the array is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from hsmc_amd import gpu

pytestmark = pytest.mark.gpu


def _flags(fn, ctx, r, sf):
    s = np.ascontiguousarray(sf, dtype=np.float64)
    out = np.full(len(s), 7, dtype=np.uint8)
    st = fn(ctx.h, r, s.ctypes.data_as(C.POINTER(C.c_double)), len(s),
            out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return st, out


def test_noalloc_scan_matches_allocating_scan():
    L = gpu.load()
    big, box_b = gpu.lattice(2, 20, 0.85)     # N = 32000
    small, box_s = gpu.lattice(2, 8, 0.90)    # N = 2048
    ctx = gpu.Context([dict(n=len(big), box=box_b, dr_max=0.1, seed=3),
                       dict(n=len(small), box=box_s, dr_max=0.1, seed=4)])
    ctx.upload(big, 0)
    ctx.upload(small, 1)
    ctx.sweeps(60)
    sf_big = np.array([1.0] + [(1.0 - k * 1e-4) ** (1.0 / 3.0) for k in range(1, 21)]
                      + [0.9995, 0.999, 0.995, 0.99])
    sf_small = np.array([1.0, 0.9999, 0.999, 0.99, 0.95])
    calls = [(0, sf_big), (1, sf_small), (0, sf_big), (1, sf_small[:2])]
    for r, sf in calls:
        st_n, got = _flags(L.hsmc_gpu_any_overlap_scaled_noalloc, ctx, r, sf)
        st_a, ref = _flags(L.hsmc_gpu_any_overlap_scaled, ctx, r, sf)
        assert st_n == 0 and st_a == 0, (st_n, st_a)
        assert np.array_equal(got, ref), (r, got, ref)
        assert set(got.tolist()) <= {0, 1}
    # the comparison covers both outcomes
    assert _flags(L.hsmc_gpu_any_overlap_scaled_noalloc, ctx, 0, sf_big)[1].min() == 0
    assert _flags(L.hsmc_gpu_any_overlap_scaled_noalloc, ctx, 0, sf_big)[1].max() == 1
    ctx.close()


def test_noalloc_scan_refuses_a_replica_beyond_its_array():
    L = gpu.load()
    conf, box = gpu.lattice(2, 60, 0.85)      # N = 864000: about 28 MB of grid
    ctx = gpu.Context([dict(n=len(conf), box=box, dr_max=0.1, seed=5)])
    ctx.upload(conf)
    st, out = _flags(L.hsmc_gpu_any_overlap_scaled_noalloc, ctx, 0, [1.0, 0.999])
    assert st == -8   # HSMC_ERR_NOT_SUPPORTED
    assert np.all(out == 7)   # nothing written
    msg = L.hsmc_gpu_last_error(ctx.h).decode()
    assert "16 MiB" in msg
    # the allocating entry point still serves it
    st, out = _flags(L.hsmc_gpu_any_overlap_scaled, ctx, 0, [1.0, 0.999])
    assert st == 0 and out[0] == 1
    ctx.close()
