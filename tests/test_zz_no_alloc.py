"""cgs_exclusive_scan_u32_noalloc through the C ABI against the allocating
cgs_exclusive_scan_u32 (gpu_exclusive_scan_total, scan.rs:312-333) on the
same inputs: outputs and totals bit-exact, on two calls in a row (the second
smaller, so stale scratch from the first would show), asynchronously without
a total, for n = 0, and a size above the scratch's limit rejected on the host."""
import ctypes as C

import numpy as np
import pytest

from cindergs_amd import _lib

pytestmark = pytest.mark.gpu


def _reference(x):
    exp = np.zeros(x.size, np.uint32)
    if x.size:
        exp[1:] = np.cumsum(x, dtype=np.uint64)[:-1].astype(np.uint32)
    return exp, int(x.sum(dtype=np.uint64)) & 0xFFFFFFFF


def test_noalloc_scan_matches_the_allocating_scan(gpu):
    rng = np.random.default_rng(5)
    for n in (1_000_003, 4_097):
        print(f"[case] noalloc scan n={n}", flush=True)
        x = rng.integers(0, 1 << 12, n, dtype=np.uint32)
        din = gpu.to_device(x)
        d_alloc = gpu.to_device(np.zeros(n, np.uint32))
        d_none = gpu.to_device(np.full(n, 0xDEADBEEF, np.uint32))
        t_alloc = gpu.exclusive_scan_total(din, d_alloc, n)
        t_none = gpu.exclusive_scan_total_noalloc(din, d_none, n)
        exp, tot = _reference(x)
        assert t_none == t_alloc == tot
        np.testing.assert_array_equal(d_none.numpy(), d_alloc.numpy())
        np.testing.assert_array_equal(d_none.numpy(), exp)
        for b in (din, d_alloc, d_none):
            b.free()


def test_noalloc_scan_without_total_and_empty(gpu):
    rng = np.random.default_rng(6)
    x = rng.integers(0, 100, 70_000, dtype=np.uint32)
    din = gpu.to_device(x)
    dout = gpu.to_device(np.zeros(x.size, np.uint32))
    assert gpu.exclusive_scan_total_noalloc(din, dout, x.size, want_total=False) is None
    gpu.synchronize()
    np.testing.assert_array_equal(dout.numpy(), _reference(x)[0])
    assert gpu.exclusive_scan_total_noalloc(None, None, 0) == 0
    din.free()
    dout.free()


def test_noalloc_scan_rejects_sizes_beyond_its_scratch(gpu):
    t = C.c_uint32()
    lib = _lib.load()
    # the size test runs before any device work: dangling pointers are never read
    rc = lib.cgs_exclusive_scan_u32_noalloc(gpu._h, C.c_void_p(16), C.c_void_p(16),
                                            (1 << 28) + 1, C.byref(t))
    assert rc == _lib.CGS_ERR_INVALID_ARG
    assert b"CGS_SCAN_NOALLOC_MAX_N" in lib.cgs_last_error()
