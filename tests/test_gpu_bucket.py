"""Stage 3 in bucket mode (per-tile buckets and a segmented sort) against the
oracle and against the LSD radix sort, on the shapes that exercise its
special paths.

Bar: counts, offsets, sorted keys, sorted ids and tile ranges bit-exact
against the oracle (the stable radix order of gpu_radix_sort_u32_pairs,
radix_sort.rs:292-368), RGBA8 within 1/255, and the LSD path
(tile_sort = 0) giving the identical frame.  Cases: tiles of tens of
thousands of entries (chunk sorts and several merge rounds, more than a
fresh context plans, so the frame re-runs Stage 3), thousands of tiles of one
or two entries (many tiles per sort window), a 3840x2160 frame (the large
LDS tile histogram) and a frame that overflows the entry arena.
"""
import numpy as np
import pytest

from cindergs_amd import synthetic as syn
from cindergs_amd._lib import Tuning
from cindergs_amd.pipeline import GaussianSceneSoa, GpuPipelineState

pytestmark = pytest.mark.gpu


def _scene_like(base, pos_scale=1.0, cov_scale=1.0):
    return GaussianSceneSoa(np.ascontiguousarray(base.positions * np.float32(pos_scale)),
                            np.ascontiguousarray(base.covariances_3d * np.float32(cov_scale)),
                            base.opacities, base.sh_coeffs)


def _check(scene, cam, comp, label, min_reruns=0):
    import oracle

    W, H = comp.screen_width, comp.screen_height
    T = comp.tiles_x * comp.tiles_y
    print(f"[case] bucket {label}: n={scene.count} {W}x{H} T={T}", flush=True)
    ref = oracle.render_frame(scene, cam, comp)
    out = {}
    for sort in (1, 0):
        with GpuPipelineState(0) as st:
            st.set_tuning(Tuning(16384, 8, 256, 1, sort))
            st.set_capture(True)
            t = st.execute_frame_gpu(scene, None, cam, comp, force_stage3=True)
            s3 = st.read_stage3(scene.count, T, with_counts=True)
            img = st.readback_framebuffer_rgba8(W, H)
        assert s3["n"] == ref["total"] == t.total_tile_entries, label
        for k, rk in (("counts", "counts"), ("offsets", "offsets"), ("keys", "sorted_keys"),
                      ("ids", "sorted_ids"), ("ranges", "ranges")):
            np.testing.assert_array_equal(s3[k], ref[rk], err_msg=f"{label} sort={sort}: {k}")
        d = int(np.abs(img.astype(np.int16) - ref["rgba"].astype(np.int16)).max())
        assert d <= 1, f"{label} sort={sort}: RGBA differs by {d}/255"
        if sort == 1:
            assert t.stage3_reruns >= min_reruns, (label, t.stage3_reruns)
        out[sort] = img
    assert np.array_equal(out[0], out[1]), label
    r = ref["ranges"]
    n = (r[:, 1] - r[:, 0]).astype(np.int64)
    print(f"[case] bucket {label}: E={ref['total']} max tile {n.max()}", flush=True)
    return n


def test_dense_tiles_take_chunk_sorts_and_merge_rounds():
    base = syn.scene_gaussian_ball(60_000, k=51)
    scene = _scene_like(base, pos_scale=0.05)
    cam, comp = syn.view_config1()
    # a fresh context plans 3 merge rounds (tiles up to 32768 entries): a
    # larger tile makes the frame run Stage 3 again with its own plan
    n = _check(scene, cam, comp, "dense 60k", min_reruns=1)
    assert n.max() > 8 * 4096


def test_many_tiny_tiles_per_window():
    base = syn.scene_gaussian_ball(30_000, k=52)
    scene = _scene_like(base, pos_scale=4.0, cov_scale=0.25)
    cam, comp = syn.view_orbit(5)
    n = _check(scene, cam, comp, "sparse 30k 1080p")
    assert np.sum((n > 0) & (n <= 2)) > 1000


def test_3840x2160_uses_the_large_histogram():
    scene = syn.scene_clusters(300_000, k=53)
    cam, comp = syn.view_orbit(11, 3840, 2160)
    assert 8192 < comp.tiles_x * comp.tiles_y <= 32768
    _check(scene, cam, comp, "300k 2160p")


def test_entry_arena_overflow_in_bucket_mode():
    base = syn.scene_gaussian_ball(20_000, k=54)
    scene = _scene_like(base, cov_scale=100.0)
    cam = syn.camera_look_at((0.0, 0.0, -3.0), (0, 0, 0), (0, 1, 0), 60.0, 1920, 1080)
    comp = syn.compositor(1920, 1080, (0.0, 0.0, -3.0))
    # the first arena holds max(2^20, 4 x splats) entries
    n = _check(scene, cam, comp, "overflow 20k", min_reruns=1)
    assert n.sum() > (1 << 20)
