"""Launch tuning (cgs_tuning, cgs_autotune) on the GPU.

Bar: every tuning candidate (the per-tile bucket sort, or the LSD radix sort
with partitions of 16384/12288/10240/8192/4096 keys and 8- or 4-bit digits;
compositor batch 512/256/128/64, heaviest-tiles-first order on/off; the presort
by depth followed by a sort of the tile bits) of a tile size gives
bit-identical tile counts, scan offsets (the presort's scan runs in depth
order and is not compared), sorted keys, sorted ids, tile ranges and RGBA8 to
the first candidate of that size (16 x 16: to the
default setting, which the parity tests hold to the oracle; the other sizes
are held to the oracle at their size in test_gpu_tile_shapes.py); the 64-bit
key path (8K) likewise for the partition sizes; the autotuner sweeps the
current tile size only, keeps one of its candidates and the frame it renders
is unchanged.
"""
import numpy as np
import pytest

from cindergs_amd import synthetic as syn
from cindergs_amd._lib import StageTimings, Tuning
from cindergs_amd.pipeline import FrameParams, tuning_candidate, tuning_candidates_count

pytestmark = pytest.mark.gpu


def _frame(gpu, cam, comp, n):
    p = FrameParams(0.0, 0.0, 1.0 / 255.0, 0.5, 0, 1)
    t = StageTimings()
    gpu.set_capture(True)
    gpu.execute_frame_resident(cam, comp, p, t)
    gpu.set_capture(False)
    W, H = comp.screen_width, comp.screen_height
    tw, th = gpu.get_tuning().tile_w, gpu.get_tuning().tile_h
    T = ((W + tw - 1) // tw) * ((H + th - 1) // th)
    s3 = gpu.read_stage3(n, T, with_counts=True)
    return s3, gpu.readback_framebuffer_rgba8(W, H)


def _same(a, b, label, offsets=True):
    """offsets=False: b's scan ran in depth order (presort), so only its per-splat
    counts, keys, ids, ranges and pixels are the same"""
    (s3a, ia), (s3b, ib) = a, b
    assert s3a["n"] == s3b["n"], f"{label}: entries"
    for k in ("counts", "offsets", "keys", "ids", "ranges"):
        if k == "offsets" and not offsets:
            continue
        np.testing.assert_array_equal(s3a[k], s3b[k], err_msg=f"{label}: {k}")
    assert np.array_equal(ia, ib), f"{label}: RGBA8"


def test_every_candidate_gives_the_same_frame(gpu):
    scene = syn.scene_clusters(150_000, k=41)
    gpu.sync_scene_and_cages(scene, None)
    cam, comp = syn.view_orbit(7, 1280, 720)
    default = gpu.get_tuning()
    assert default.as_dict() == {"radix_partition_keys": 16384, "radix_digit_bits": 8,
                                 "composite_batch": 64, "tile_order": 1, "tile_sort": 0,
                                 "tile_w": 16, "tile_h": 16}
    refs = {(16, 16): _frame(gpu, cam, comp, scene.count)}
    n = tuning_candidates_count()
    seen = set()
    try:
        for i in range(n):
            c = tuning_candidate(i)
            seen.add(tuple(c.as_dict().values()))
            print(f"[case] candidate {i}: {c.as_dict()}", flush=True)
            gpu.set_tuning(c)
            got = _frame(gpu, cam, comp, scene.count)
            size = (c.tile_w, c.tile_h)
            if size in refs:
                _same(refs[size], got, f"candidate {i}", offsets=c.tile_sort != 2)
            else:
                assert c.tile_sort != 2
                refs[size] = got
    finally:
        gpu.set_tuning(default)
    assert len(seen) == n == 120
    assert set(refs) == {(16, 16), (8, 8), (16, 32), (32, 16)}


def test_partition_sizes_on_64bit_keys(gpu):
    scene = syn.scene_gaussian_ball(40_000, k=42)
    gpu.sync_scene_and_cages(scene, None)
    cam = syn.camera_look_at((0.0, 0.0, -3.0), (0, 0, 0), (0, 1, 0), 60.0, 7680, 4320)
    comp = syn.compositor(7680, 4320, (0.0, 0.0, -3.0))
    default = gpu.get_tuning()
    ref = _frame(gpu, cam, comp, scene.count)
    for pk, batch in ((8192, 128), (10240, 256), (16384, 64), (4096, 256), (12288, 128)):
        t = Tuning(pk, 8, batch, 0)
        print(f"[case] 8K partition {pk} batch {batch}", flush=True)
        gpu.set_tuning(t)
        _same(ref, _frame(gpu, cam, comp, scene.count), f"8K partition {pk}")
    gpu.set_tuning(default)


def test_autotune_keeps_a_candidate_and_the_frame(gpu):
    scene = syn.scene_clusters(120_000, k=43)
    gpu.sync_scene_and_cages(scene, None)
    cam, comp = syn.view_orbit(3, 960, 540)
    p = FrameParams(0.0, 0.0, 1.0 / 255.0, 0.5, 0, 0)
    default = gpu.get_tuning()
    ref = _frame(gpu, cam, comp, scene.count)
    best, table = gpu.autotune(cam, comp, p, reps=2)
    print(f"[case] autotune best {best.as_dict()}", flush=True)
    assert any(c.as_dict() == best.as_dict() for c, _ in table)
    assert (best.tile_w, best.tile_h) == (16, 16)
    assert all((ms > 0) == ((c.tile_w, c.tile_h) == (16, 16)) for c, ms in table)
    assert all(ms == -1.0 for c, ms in table if (c.tile_w, c.tile_h) != (16, 16))
    assert gpu.get_tuning().as_dict() == best.as_dict()
    _same(ref, _frame(gpu, cam, comp, scene.count), "autotuned")
    gpu.set_tuning(default)


def test_invalid_tuning_is_rejected(gpu):
    from cindergs_amd._lib import GpuError

    for bad in (Tuning(2048, 8, 256, 1), Tuning(12000, 8, 256, 1), Tuning(16384, 6, 256, 1), Tuning(16384, 8, 100, 1),
                Tuning(16384, 8, 256, 2), Tuning(16384, 8, 256, 1, 3),
                Tuning(16384, 8, 256, 1, 0, 32, 32), Tuning(16384, 8, 256, 1, 0, 8, 16),
                Tuning(16384, 8, 256, 1, 0, 0, 0), Tuning(16384, 8, 128, 1, 0, 8, 8),
                Tuning(16384, 8, 512, 1, 0, 8, 8), Tuning(16384, 8, 512, 1, 0, 16, 32),
                Tuning(16384, 8, 1024, 1)):
        with pytest.raises(GpuError):
            gpu.set_tuning(bad)
