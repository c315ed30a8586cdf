"""The scan entry point that allocates nothing (pg_scan_u32_static, whose
workspace is a static device array per GPU) against pg_scan_u32 with a
caller-allocated workspace: bit-exact inclusive scans and totals, two calls
in a row with the second smaller, with and without caller tile sums, and the
refusal of a size beyond the static workspace."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan(gpu, host, static, tile_sums=None):
    n = host.size
    d = gpu.DeviceArray.from_numpy(host)
    total = ctypes.c_uint64()
    ts = gpu.DeviceArray.from_numpy(tile_sums) if tile_sums is not None else None
    tsp = ts.ptr if ts is not None else None
    if static:
        rc = gpu.lib().pg_scan_u32_static(d.ptr, n, tsp, ctypes.byref(total), None)
    else:
        tiles = (n + 255) // 256
        scratch = gpu.DeviceArray(8 * (tiles + 4) + 4 * tiles)
        rc = gpu.lib().pg_scan_u32(d.ptr, n, tsp, scratch.ptr, ctypes.byref(total), None)
    assert rc == 0, gpu.last_error()
    gpu.device_synchronize()
    return d.to_numpy(np.uint32, n), total.value


def _tile_sums(host):
    tiles = (host.size + 255) // 256
    pad = np.zeros(tiles * 256 - host.size, dtype=np.uint32)
    return np.concatenate([host, pad]).reshape(tiles, 256).sum(axis=1, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("sizes", [(9_000_001, 1_000_003), (67_110_147, 300), (2_000_000, 1)])
def test_static_scan_equals_workspace_scan(gpu, sizes):
    # two calls in a row on the shared static workspace, the second smaller:
    # nothing the first call left there may leak into the second
    rng = np.random.default_rng(sum(sizes))
    for i, n in enumerate(sizes):
        host = rng.integers(0, 219, size=n, dtype=np.uint32)
        print(f"[gpu] static scan call {i}: n={n}", flush=True)
        want, want_total = _scan(gpu, host, static=False)
        got, got_total = _scan(gpu, host, static=True)
        assert np.array_equal(got, want) and got_total == want_total
        assert got_total == int(host.sum(dtype=np.uint64))
        ts = _tile_sums(host)
        got2, tot2 = _scan(gpu, host, static=True, tile_sums=ts)
        assert np.array_equal(got2, want) and tot2 == want_total


def test_static_scan_at_its_maximum_and_beyond(gpu):
    n = 1 << 28
    host = np.ones(n, dtype=np.uint32)
    print(f"[gpu] static scan at the maximum n={n}", flush=True)
    got, total = _scan(gpu, host, static=True)
    assert total == n and got[0] == 1 and got[-1] == n and got[n // 2] == n // 2 + 1
    d = gpu.DeviceArray(4)
    t = ctypes.c_uint64()
    assert gpu.lib().pg_scan_u32_static(d.ptr, n + 1, None, ctypes.byref(t), None) == gpu.PG_ERR_INVALID
    assert "exceeds" in gpu.last_error()
    # a call after the refusal still scans correctly
    host = np.arange(1000, dtype=np.uint32)
    got, total = _scan(gpu, host, static=True)
    assert np.array_equal(got, np.cumsum(host, dtype=np.uint64).astype(np.uint32)) and total == int(host.sum())
