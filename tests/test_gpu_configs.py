"""The workloads at their stated scales, against the oracle.

* a 3840x2160 view of the 4,000,000-splat scene rendered inside a view batch
  (cgs_execute_views_batched) together with other views of the same scene:
  the batch's pixels equal the per-view frame bit for bit, whose tile counts,
  keys, ids and ranges equal the oracle's exactly; the batch's RGBA is within
  1/255 of the oracle's;
* the f16 storage variant (project_gaussians_v2_f16_kernel, project.rs:
  1344-1679) on the full 1,073,930-splat scene at 1920x1080: projected
  attributes bitwise, keys / ids / ranges exact, RGBA within 1/255;
* the 7680x4320 frame of the full 1,073,930-splat scene (129,600 tiles, the
  64-bit keys of emit_tile_keys_u64_kernel and gpu_radix_sort_u64_pairs,
  binning.rs:254-328, radix_sort.rs:377-477): keys / ids / ranges exact,
  RGBA within 1/255.
"""
import numpy as np
import pytest

from cindergs_amd import synthetic as syn
from cindergs_amd._lib import StageTimings
from cindergs_amd.pipeline import FrameParams, GpuPipelineState, ViewBatch

pytestmark = pytest.mark.gpu


def _params():
    return FrameParams(0.0, 0.0, 1.0 / 255.0, 0.5, 0, 1)


def _stage3_vs_oracle(st, ref, n, T, label):
    s3 = st.read_stage3(n, T)
    assert s3["n"] == ref["total"], label
    np.testing.assert_array_equal(s3["keys"], ref["sorted_keys"], err_msg=f"{label}: keys")
    np.testing.assert_array_equal(s3["ids"], ref["sorted_ids"], err_msg=f"{label}: ids")
    np.testing.assert_array_equal(s3["ranges"], ref["ranges"], err_msg=f"{label}: ranges")


def _rgba_close(img, ref, label):
    d = int(np.abs(img.astype(np.int16) - ref["rgba"].astype(np.int16)).max())
    print(f"[case] {label}: E={ref['total']} max|dRGBA|={d}", flush=True)
    assert d <= 1, f"{label}: RGBA differs by {d}/255"


def test_2160p_view_of_the_4m_scene_in_a_batch(oracle_lib):
    import oracle

    scene = syn.scene_gaussian_ball(4_000_000, k=3)
    view_scene, cams, comps = syn.batch_views(64, 3)
    # the batch workload's 3840x2160 view 2 of scene 2 (4M splats), with two
    # lower-resolution views of the same scene
    pick = [2] + [v for v in range(64) if view_scene[v] == 2 and comps[v].screen_width < 3840][:2]
    assert comps[2].screen_width == 3840 and comps[2].screen_height == 2160
    cams = [cams[v] for v in pick]
    comps = [comps[v] for v in pick]
    with GpuPipelineState(0) as st:
        st.sync_scene_and_cages(scene, None)
        with ViewBatch(0, lanes=2) as batch:
            stats = batch.execute([st], [0] * len(pick), cams, comps, _params())
            print(f"[case] batch of views {pick}: {stats.as_dict()}", flush=True)
            got = [batch.readback_rgba8(v) for v in range(len(pick))]
            counts = [batch.view_counts(v) for v in range(len(pick))]
        for v in range(len(pick)):
            W, H = comps[v].screen_width, comps[v].screen_height
            t = StageTimings()
            st.execute_frame_resident(cams[v], comps[v], _params(), t)
            own = st.readback_framebuffer_rgba8(W, H)
            assert np.array_equal(got[v], own), f"view {pick[v]}: batch differs from the frame"
            assert counts[v] == (t.total_tile_entries, t.surviving_gaussians)
            if v == 0:
                ref = oracle.render_frame(scene, cams[v], comps[v])
                _stage3_vs_oracle(st, ref, scene.count, comps[v].tiles_x * comps[v].tiles_y,
                                  "4M 2160p")
                _rgba_close(got[v], ref, "4M 2160p batched")


def test_f16_storage_on_the_full_scene(oracle_lib):
    import oracle

    scene = syn.scene_clusters()
    q = syn.quantize_f16(scene, 3)
    cam, comp = syn.view_orbit(17)
    cov_h = q.covariances_f16_packed.view(np.uint16).reshape(-1, 6).copy()
    sh_h = q.sh_coeffs_f16_packed.view(np.uint16).reshape(q.count, -1).copy()
    with GpuPipelineState(0) as st:
        t = st.execute_frame_gpu(q, None, cam, comp, force_stage3=True)
        ref = oracle.render_frame(scene, cam, comp, f16=(cov_h, sh_h))
        attrs, ext = st.read_projected(scene.count)
        np.testing.assert_array_equal(ext, ref["extents"])
        alive = ext != 0
        np.testing.assert_array_equal(attrs[alive].view(np.uint32),
                                      ref["attrs"][alive].view(np.uint32))
        assert t.total_tile_entries == ref["total"]
        _stage3_vs_oracle(st, ref, scene.count, comp.tiles_x * comp.tiles_y, "f16 1.07M")
        _rgba_close(st.readback_framebuffer_rgba8(1920, 1080), ref, "f16 1.07M 1080p")


def test_7680x4320_frame_of_the_full_scene(oracle_lib):
    import oracle

    scene = syn.scene_clusters()
    cam, comp = syn.view_orbit(3, 7680, 4320)
    assert comp.tiles_x * comp.tiles_y == 129_600
    with GpuPipelineState(0) as st:
        t = st.execute_frame_gpu(scene, None, cam, comp, force_stage3=True)
        ref = oracle.render_frame(scene, cam, comp)
        assert t.total_tile_entries == ref["total"]
        _stage3_vs_oracle(st, ref, scene.count, 129_600, "8K 1.07M")
        _rgba_close(st.readback_framebuffer_rgba8(7680, 4320), ref, "8K 1.07M")
