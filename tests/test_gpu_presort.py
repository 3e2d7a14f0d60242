"""Stage 3 with the splats presorted by depth (cgs_tuning.tile_sort = 2)
against the oracle.

The mode sorts the splats once by their key's depth bits (stable, so ties
keep splat order), bins them in that order, so that the entries come out in
(depth, splat) order, and sorts the entries by the tile bits only: 2 passes
over the entries at 1080p instead of 4, 3 instead of 7 on 64-bit keys.

Bar (the parity tests' bar): per-splat tile counts, sorted keys, sorted ids
and tile ranges bit-exact against the oracle (the reference's stable order of
gpu_radix_sort_u32_pairs / _u64_pairs, radix_sort.rs:292-477), RGBA8 within
1/255, and the frame identical to the LSD path's (tile_sort = 0): the
1.07M-splat 1080p orbit, an 8K frame (64-bit keys), the 3D mip filter, a
cage-deformed scene with split children, 8 x 8 tiles, an arena overflow, and
a view batch.
"""
import numpy as np
import pytest

from cindergs_amd import synthetic as syn
from cindergs_amd._lib import StageTimings, Tuning
from cindergs_amd.pipeline import FrameParams, GpuPipelineState, ViewBatch

pytestmark = pytest.mark.gpu


def _tuning(sort, tile=(16, 16), part=16384):
    return Tuning(part, 8, 64 if tile == (8, 8) else 256, 1, sort, tile[0], tile[1])


def _check(scene, cam, comp, label, tile=(16, 16), mip=0.0, part=16384):
    import oracle

    W, H = comp.screen_width, comp.screen_height
    T = ((W + tile[0] - 1) // tile[0]) * ((H + tile[1] - 1) // tile[1])
    print(f"[case] presort {label}: n={scene.count} {W}x{H} T={T}", flush=True)
    ref = oracle.render_frame(scene, cam, comp, mip=mip, tile=tile)
    imgs = {}
    for sort in (2, 0):
        with GpuPipelineState(0) as st:
            st.set_tuning(_tuning(sort, tile, part))
            st.set_capture(True)
            t = st.execute_frame_gpu(scene, None, cam, comp, mip_filter_3d_var=mip,
                                     force_stage3=True)
            s3 = st.read_stage3(scene.count, T, with_counts=True)
            imgs[sort] = st.readback_framebuffer_rgba8(W, H)
        assert s3["n"] == ref["total"] == t.total_tile_entries, (label, sort)
        for k, rk in (("counts", "counts"), ("keys", "sorted_keys"), ("ids", "sorted_ids"),
                      ("ranges", "ranges")):
            np.testing.assert_array_equal(s3[k], ref[rk], err_msg=f"{label} sort={sort}: {k}")
        if sort == 0:
            np.testing.assert_array_equal(s3["offsets"], ref["offsets"], err_msg=label)
        d = int(np.abs(imgs[sort].astype(np.int16) - ref["rgba"].astype(np.int16)).max())
        print(f"[case] presort {label} sort={sort}: E={s3['n']} max|dRGBA|={d}", flush=True)
        assert d <= 1, f"{label} sort={sort}: RGBA differs by {d}/255"
    assert np.array_equal(imgs[2], imgs[0]), f"{label}: presort and LSD frames differ"


def test_presort_1080p_orbit(oracle_lib):
    scene = syn.scene_clusters()
    for frame in (0, 53):
        cam, comp = syn.view_orbit(frame)
        _check(scene, cam, comp, f"1.07M 1080p frame {frame}")


@pytest.mark.parametrize("part", [10240, 8192])
def test_presort_partition_sizes(oracle_lib, part):
    scene = syn.scene_gaussian_ball(300_000, k=81)
    cam, comp = syn.view_orbit(9, 1280, 720)
    _check(scene, cam, comp, f"300k 720p partitions {part}", part=part)


def test_presort_8k_64bit_keys(oracle_lib):
    scene = syn.scene_gaussian_ball(200_000, k=82)
    eye = (0.0, 0.0, -3.0)
    cam = syn.camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, 7680, 4320)
    comp = syn.compositor(7680, 4320, eye)
    _check(scene, cam, comp, "200k 8K (64-bit keys)")


def test_presort_8x8_tiles(oracle_lib):
    scene = syn.scene_clusters(150_000, k=83)
    cam, comp = syn.view_orbit(17, 1001, 557)
    _check(scene, cam, comp, "150k 1001x557 8x8", tile=(8, 8))


def test_presort_mip_filter(oracle_lib):
    scene = syn.scene_gaussian_ball(100_000, k=84)
    cam, comp = syn.view_config1()
    _check(scene, cam, comp, "100k 800x600 mip 2e-3", mip=2e-3)


def test_presort_small_and_empty_frames(oracle_lib):
    scene = syn.scene_gaussian_ball(3_000, k=85)
    for W, H in ((1, 1), (17, 9), (64, 48)):
        cam = syn.camera_look_at((0.0, 0.0, -3.0), (0, 0, 0), (0, 1, 0), 60.0, W, H)
        _check(scene, cam, syn.compositor(W, H, (0.0, 0.0, -3.0)), f"3k {W}x{H}")


def test_presort_deformed_scene_with_splits():
    """Stage 1 with splitting (children appended after the rest splats): the
    presort covers the children; the frame equals the LSD path's."""
    from cindergs_amd.cages import TetMesh

    scene = syn.scene_gaussian_ball(40_000, k=86)
    lo, hi = scene.positions.min(0), scene.positions.max(0)
    mesh = TetMesh.jittered_grid(lo, hi, (8, 8, 8))
    tet, w = mesh.bind_gaussians(scene.positions)
    moved = (mesh.rest_vertices * np.array([1.6, 1.0, 1.0], np.float32)).astype(np.float32)
    cage = mesh.cage(moved, tet, w)
    cam, comp = syn.view_config1()
    p = FrameParams(1.02, 0.0, 1.0 / 255.0, 0.5, 0, 1)
    out = {}
    for sort in (2, 0):
        with GpuPipelineState(0) as st:
            st.sync_scene_and_cages(scene, cage)
            st.set_tuning(_tuning(sort))
            t = StageTimings()
            st.execute_frame_resident(cam, comp, p, t)
            s3 = st.read_stage3(2 * scene.count, comp.tiles_x * comp.tiles_y)
            out[sort] = (int(t.total_tile_entries), int(t.active_gaussians), s3["keys"].copy(),
                         s3["ids"].copy(), s3["ranges"].copy(),
                         st.readback_framebuffer_rgba8(800, 600))
    print(f"[case] presort deformed: E={out[2][0]} active={out[2][1]}", flush=True)
    assert out[2][1] > scene.count, "no splat was split"
    assert out[2][:2] == out[0][:2]
    for k in (2, 3, 4, 5):
        assert np.array_equal(out[2][k], out[0][k]), k


def test_presort_entry_arena_overflow(oracle_lib):
    """Splats 10x wider: more entries than a fresh context's arena holds, so
    the frame re-runs Stage 3 with a larger arena."""
    from cindergs_amd.pipeline import GaussianSceneSoa

    base = syn.scene_gaussian_ball(20_000, k=89)
    scene = GaussianSceneSoa(base.positions, np.ascontiguousarray(base.covariances_3d * 100.0),
                             base.opacities, base.sh_coeffs)
    cam = syn.camera_look_at((0.0, 0.0, -3.0), (0, 0, 0), (0, 1, 0), 60.0, 1920, 1080)
    _check(scene, cam, syn.compositor(1920, 1080, (0.0, 0.0, -3.0)), "overflow 20k")


def test_presort_batch_matches_per_view_frames():
    scenes = [syn.scene_gaussian_ball(30_000, k=87), syn.scene_clusters(60_000, k=88)]
    tune = _tuning(2)
    states = []
    for sc in scenes:
        st = GpuPipelineState(0)
        st.sync_scene_and_cages(sc, None)
        st.set_tuning(tune)
        states.append(st)
    view_scene, cams, comps = syn.batch_views(6, 2, seed=19,
                                              resolutions=((320, 240), (800, 600), (1920, 1080)))
    p = FrameParams(0.0, 0.0, 1.0 / 255.0, 0.5, 0, 0)
    with ViewBatch(0, lanes=3) as batch:
        batch.set_tuning(tune)
        for rep in range(2):
            stats = batch.execute(states, view_scene, cams, comps, p)
            print(f"[case] presort batch {rep}: {stats.as_dict()}", flush=True)
            for v in range(len(cams)):
                st = states[view_scene[v]]
                t = StageTimings()
                st.execute_frame_resident(cams[v], comps[v], p, t)
                W, H = comps[v].screen_width, comps[v].screen_height
                assert np.array_equal(batch.readback_rgba8(v), st.readback_framebuffer_rgba8(W, H))
                assert batch.view_counts(v)[0] == int(t.total_tile_entries)
    for st in states:
        st.close()
