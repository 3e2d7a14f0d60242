"""SortCache hysteresis inside execute_frame_gpu (cgs_set_sort_hysteresis): the
CPU fallback frame's reuse of the previous binning while the camera and the
cages move by at most 1e-5 (src/render/pipeline.rs:552-600,
execute_cpu_reference_with_options 417-507).

Bar: on a reuse, Stage 2 bit-exact against the oracle for the new camera,
Stage 3 outputs identical to the previous frame's, and the image within
1/255 of the oracle compositor run on the new projection in the previous
tile order (splats culled now: zero records); above the bound, or after
invalidate_stage3, the frame equals a fresh one.
"""
import ctypes as C
import math

import numpy as np
import pytest

from cindergs_amd import _lib
from cindergs_amd import synthetic as syn
from cindergs_amd.cages import CageSpringSimulator, TetMesh

pytestmark = pytest.mark.gpu


def _moved(cam, k, delta):
    c = _lib.CameraUniforms.from_buffer_copy(cam)
    c.view_matrix[k] = np.float32(c.view_matrix[k] + np.float32(delta))
    return c


def _expected(oracle, attrs, ext, ids, ranges, comp):
    a = attrs.copy()
    a[ext == 0] = 0.0
    return oracle.composite(a, ids, ranges, comp.screen_width, comp.screen_height,
                            list(comp.background_color))


def test_camera_jitter_reuses_the_tile_order(gpu, oracle_lib):
    import oracle

    scene = syn.scene_gaussian_ball(30_000, k=51)
    cam, comp = syn.view_config1()
    tiles = comp.tiles_x * comp.tiles_y
    n = scene.count
    print("[case] hysteresis camera", flush=True)
    gpu.set_sort_hysteresis(True)
    gpu.set_capture(True)
    try:
        t0 = gpu.execute_frame_gpu(scene, None, cam, comp)
        assert t0.stage3_cached == 0
        s3a = gpu.read_stage3(n, tiles)
        cam2 = _moved(cam, 12, 0.8e-5)
        t1 = gpu.execute_frame_gpu(scene, None, cam2, comp)
        assert t1.stage3_cached == 1 and gpu.sort_reuse_count == 1
        s3b = gpu.read_stage3(n, tiles)
        for k in ("keys", "ids", "ranges"):
            np.testing.assert_array_equal(s3b[k], s3a[k])
        ref = oracle.render_frame(scene, cam2, comp)
        attrs, ext = gpu.read_projected(n)
        np.testing.assert_array_equal(ext, ref["extents"])
        alive = ext != 0
        np.testing.assert_array_equal(attrs[alive].view(np.uint32),
                                      ref["attrs"][alive].view(np.uint32))
        exp = _expected(oracle, ref["attrs"], ref["extents"], s3a["ids"], s3a["ranges"], comp)
        img = gpu.readback_framebuffer_rgba8(800, 600)
        assert np.abs(img.astype(np.int16) - exp.astype(np.int16)).max() <= 1
        # re-anchored: another 0.8e-5 from cam2 reuses again
        cam3 = _moved(cam2, 12, 0.8e-5)
        assert gpu.execute_frame_gpu(scene, None, cam3, comp).stage3_cached == 1
        # above the bound: a fresh binning, equal to the oracle's
        cam4 = _moved(cam3, 12, 3e-5)
        assert gpu.execute_frame_gpu(scene, None, cam4, comp).stage3_cached == 0
        ref4 = oracle.render_frame(scene, cam4, comp)
        s3d = gpu.read_stage3(n, tiles)
        np.testing.assert_array_equal(s3d["ids"], ref4["sorted_ids"])
        np.testing.assert_array_equal(s3d["ranges"], ref4["ranges"])
        img = gpu.readback_framebuffer_rgba8(800, 600)
        assert np.abs(img.astype(np.int16) - ref4["rgba"].astype(np.int16)).max() <= 1
        # invalidate_stage3 drops the cache
        gpu.invalidate_stage3()
        assert gpu.execute_frame_gpu(scene, None, cam4, comp).stage3_cached == 0
        # a changed frame parameter (mip) never reuses
        t = gpu.execute_frame_gpu(scene, None, cam4, comp, mip_filter_3d_var=1e-4)
        assert t.stage3_cached == 0
        assert gpu.sort_reuse_count == 2
    finally:
        gpu.set_sort_hysteresis(False)
        gpu.set_capture(False)


def test_cage_jitter_reuses_the_tile_order(gpu, oracle_lib):
    """A cage vertex moved by 0.5e-5: Stage 1 re-runs, Stage 3 is reused, the
    image is the new deformed projection in the previous order."""
    import oracle

    scene = syn.scene_gaussian_ball(20_000, k=52)
    lo, hi = scene.positions.min(0), scene.positions.max(0)
    mesh = TetMesh.jittered_grid(lo, hi, (8, 8, 8))
    tet, w = mesh.bind_gaussians(scene.positions)
    sim = CageSpringSimulator(mesh.rest_vertices, stiffness=40.0, damping=math.exp(-4.0 / 60.0))
    sim.impulse(0.2, 0.05)
    for _ in range(5):
        sim.step(1.0 / 60.0)
    cage = mesh.cage(sim.x, tet, w)
    v2 = cage.vertices.copy()
    v2[:, 1] += np.float32(0.5e-5)
    cage2 = mesh.cage(v2, tet, w)
    cam, comp = syn.view_config1()
    tiles = comp.tiles_x * comp.tiles_y
    n = scene.count
    print("[case] hysteresis cage", flush=True)
    gpu.set_sort_hysteresis(True)
    gpu.set_capture(True)
    try:
        gpu.execute_frame_gpu(scene, cage, cam, comp)
        s3a = gpu.read_stage3(n, tiles)
        t1 = gpu.execute_frame_gpu(scene, cage2, cam, comp)
        assert t1.stage1_skipped == 0 and t1.stage3_cached == 1
        ref = oracle.render_frame_deformed(scene, cage2, cam, comp)
        attrs, ext = gpu.read_projected(n)
        np.testing.assert_array_equal(ext, ref["extents"][:n])
        exp = _expected(oracle, ref["attrs"][:n], ref["extents"][:n], s3a["ids"],
                        s3a["ranges"], comp)
        img = gpu.readback_framebuffer_rgba8(800, 600)
        assert np.abs(img.astype(np.int16) - exp.astype(np.int16)).max() <= 1
        # a cage move above the bound re-sorts
        v3 = v2.copy()
        v3[0, 0] += np.float32(1e-3)
        t2 = gpu.execute_frame_gpu(scene, mesh.cage(v3, tet, w), cam, comp)
        assert t2.stage3_cached == 0
    finally:
        gpu.set_sort_hysteresis(False)
        gpu.set_capture(False)


def test_exact_key_mode_is_unchanged(gpu):
    """With hysteresis off (the default) a 0.8e-5 camera move re-sorts: the
    Stage 3 cache key is the exact bits of gpu_state.rs:454-499."""
    scene = syn.scene_gaussian_ball(5_000, k=53)
    cam, comp = syn.view_config1()
    gpu.execute_frame_gpu(scene, None, cam, comp)
    assert gpu.execute_frame_gpu(scene, None, cam, comp).stage3_cached == 1
    assert gpu.execute_frame_gpu(scene, None, _moved(cam, 12, 0.8e-5), comp).stage3_cached == 0
    r = C.c_uint64()
    _lib.check(gpu._lib.cgs_sort_hysteresis_reuse_count(gpu._h, C.byref(r)))
    assert r.value == 0
