"""CPU-side checks of the C ABI library: it loads, exports every entry point
declared in include/cindergs.h, and reports the error taxonomy of the
reference (GpuError, src/render/gpu.rs:14-30) without a device."""
import ctypes as C
import os
import re

import pytest

from cindergs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "cindergs.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|uint32_t|uint64_t|const char\*)\s+(cgs_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_struct_layouts_match_reference_sizes():
    assert C.sizeof(_lib.CameraUniforms) == 96   # camera.rs:9-36
    assert C.sizeof(_lib.CompositorUniforms) == 48  # rasterizer/mod.rs:91-94


def test_invalid_arguments_fail_loudly():
    lib = _lib.load()
    assert lib.cgs_create(0, None) == _lib.CGS_ERR_INVALID_ARG
    assert b"null" in lib.cgs_last_error()
    assert lib.cgs_invalidate(None) == _lib.CGS_ERR_INVALID_ARG
    n = C.c_uint32()
    assert lib.cgs_kernel_times(None, None, None, 0, C.byref(n)) == _lib.CGS_ERR_INVALID_ARG


def test_no_device_maps_to_noadapter():
    """Without a visible GPU, cgs_create reports NoAdapter (the reference caches
    the acquisition failure, pipeline.rs:211-213, and callers fall back above
    the ABI)."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    code = _lib.load().cgs_create(0, C.byref(h))
    assert code == _lib.CGS_ERR_NO_DEVICE
    with pytest.raises(_lib.GpuError) as ei:
        _lib.check(code)
    assert ei.value.kind == "NoAdapter"


def test_tuning_candidates_cover_every_knob():
    """The autotuner's sweep (host-only entry points, no GPU needed).  16 x 16
    tiles: the bucket sort with every compositor batch and tile order, and the
    LSD radix sort with every combination of partition size, digit width,
    compositor batch and tile order.  8 x 8 tiles (batch 64): radix partitions
    16384 / 8192 and the bucket sort, each with both orders.  16 x 32 and
    32 x 16: the same with batches 256 / 128.  Then the presort (tile_sort 2):
    16 x 16 with partitions 16384 / 10240 / 8192 x batch 256 / 128 x order, and
    8 x 8 with partitions 16384 / 8192 x order.  Last, 16 x 16 with compositor
    batches of 512: partitions 16384 / 12288 / 10240 / 8192 x order.  Each
    exactly once, 16 x 16 first."""
    from cindergs_amd.pipeline import tuning_candidate, tuning_candidates_count

    n = tuning_candidates_count()
    combos = [tuple(tuning_candidate(i).as_dict().values()) for i in range(n)]
    assert n == len(set(combos)) == 66 + 6 + 12 + 12 + 16 + 8
    assert sorted(combos[112:]) == sorted(
        [(p, 8, 512, o, 0, 16, 16) for p in (16384, 12288, 10240, 8192) for o in (0, 1)])
    presort = combos[96:112]
    assert all(c[4] == 2 for c in presort) and not any(c[4] == 2 for c in combos[:96])
    assert sorted(presort) == sorted(
        [(p, 8, b, o, 2, 16, 16) for p in (16384, 10240, 8192) for b in (256, 128) for o in (0, 1)]
        + [(p, 8, 64, o, 2, 8, 8) for p in (16384, 8192) for o in (0, 1)])
    combos = combos[:96]
    base = [c for c in combos if c[5:] == (16, 16)]
    assert combos[:66] == base
    bucket = [c for c in base if c[4] == 1]
    radix = [c for c in base if c[4] == 0]
    assert len(bucket) == 6 and {(c[2], c[3]) for c in bucket} == {
        (b, o) for b in (256, 128, 64) for o in (0, 1)}
    assert {c[0] for c in radix} == {16384, 12288, 10240, 8192, 4096}
    assert {c[1] for c in radix} == {8, 4}
    assert {c[2] for c in radix} == {256, 128, 64}
    assert {c[3] for c in radix} == {0, 1}
    small = [c for c in combos if c[5:] == (8, 8)]
    assert sorted(small) == sorted(
        [(p, 8, 64, o, 0, 8, 8) for p in (16384, 8192) for o in (0, 1)]
        + [(16384, 8, 64, o, 1, 8, 8) for o in (0, 1)])
    for shape in ((16, 32), (32, 16)):
        got = [c for c in combos if c[5:] == shape]
        assert sorted(got) == sorted(
            [(p, 8, b, o, 0) + shape for p in (16384, 8192) for b in (256, 128) for o in (0, 1)]
            + [(16384, 8, b, o, 1) + shape for b in (256, 128) for o in (0, 1)])
    assert tuning_candidate(0).as_dict() == {"radix_partition_keys": 16384, "radix_digit_bits": 8,
                                             "composite_batch": 256, "tile_order": 1,
                                             "tile_sort": 0, "tile_w": 16, "tile_h": 16}
