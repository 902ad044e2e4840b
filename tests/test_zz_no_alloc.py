"""The sort entry point that allocates nothing (lucida_sort_u32_device_noalloc),
called through the C ABI and checked against lucida_sort_u32_device, the
entry point whose scratch it replaces: bit-identical keys and values on two
calls in a row (the second smaller, so stale scratch from the first would
show), plus the size limit. Each input is also checked against numpy's
stable sort, the (key, index) order the reference's radix tests expect
(crates/gpu/tests/radix_sort.rs)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import lucida_amd as L

pytestmark = pytest.mark.gpu


def _dev(gpu, lib, nbytes):
    p = C.c_void_p()
    L._check(lib.lucida_device_alloc(gpu._h, max(nbytes, 4), C.byref(p)))
    return p


def _sort_with(gpu, lib, fn, keys, vals):
    n = len(keys)
    bufs = [_dev(gpu, lib, n * 4) for _ in range(4)]
    try:
        L._check(lib.lucida_copy_to_device(gpu._h, bufs[0], keys.ctypes.data, n * 4))
        L._check(lib.lucida_copy_to_device(gpu._h, bufs[1], vals.ctypes.data, n * 4))
        L._check(fn(gpu._h, bufs[0], bufs[1], bufs[2], bufs[3], n))
        ko, vo = np.empty(n, np.uint32), np.empty(n, np.uint32)
        L._check(lib.lucida_copy_to_host(gpu._h, ko.ctypes.data, bufs[0], n * 4))
        L._check(lib.lucida_copy_to_host(gpu._h, vo.ctypes.data, bufs[1], n * 4))
        gpu.check_errors()
        return ko, vo
    finally:
        for b in bufs:
            lib.lucida_device_free(gpu._h, b)


def test_noalloc_sort_matches_the_allocating_sort_on_two_calls(gpu):
    lib = L.lib()
    rng = np.random.default_rng(7)
    # 3,000,017 keys (a ragged last tile) with many duplicate keys, then a
    # smaller second call: the shared scratch keeps the first call's
    # histograms, which the second must not read.
    for n, hi in ((3_000_017, 1 << 20), (100_003, 1 << 32)):
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        print(f"noalloc sort: n={n}", flush=True)
        k_ref, v_ref = _sort_with(gpu, lib, lib.lucida_sort_u32_device, keys, vals)
        k_na, v_na = _sort_with(gpu, lib, lib.lucida_sort_u32_device_noalloc, keys, vals)
        assert np.array_equal(k_na, k_ref) and np.array_equal(v_na, v_ref)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k_na, keys[order]) and np.array_equal(v_na, vals[order])


def test_noalloc_sort_at_its_limit_and_beyond(gpu):
    lib = L.lib()
    n = int(L.SORT_NOALLOC_MAX_KEYS)
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    print(f"noalloc sort: n={n} (the limit)", flush=True)
    k_na, v_na = _sort_with(gpu, lib, lib.lucida_sort_u32_device_noalloc, keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k_na, keys[order]) and np.array_equal(v_na, vals[order])
    # One key more is refused before anything is queued (the pointers are
    # never dereferenced on this path).
    p = C.c_void_p(16)
    st = lib.lucida_sort_u32_device_noalloc(gpu._h, p, p, p, p, n + 1)
    assert st == 6, st  # LUCIDA_INVALID_ARG
    assert b"LUCIDA_SORT_NOALLOC_MAX_KEYS" in lib.lucida_last_error()
    # n == 0 is not an error
    assert lib.lucida_sort_u32_device_noalloc(gpu._h, None, None, None, None, 0) == 0
