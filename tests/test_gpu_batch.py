"""Batched multi-view rendering (cgs_execute_views_batched) on the GPU.

Bar: every view of a batch is bit-identical (RGBA8, tile-entry and survivor
counts) to cgs_execute_frame on the same scene, camera and parameters, which
the parity tests hold to the oracle; one batched view is also checked against
the oracle directly (keys exact there, RGBA within 1/255).  Covers several
scenes and resolutions in one batch, an empty scene (background frame), a
cage-deformed scene with splitting, the 64-bit key path (8K view), graph
reuse on an unchanged batch and re-capture after a change.
"""
import math

import numpy as np
import pytest

from cindergs_amd import synthetic as syn
from cindergs_amd._lib import StageTimings
from cindergs_amd.pipeline import (FrameParams, GaussianSceneSoa, GpuPipelineState,
                                   ViewBatch)

pytestmark = pytest.mark.gpu

SMALL = ((64, 48), (320, 240), (200, 150), (800, 600))


def _params(tau=0.0, mip=0.0):
    return FrameParams(tau, mip, 1.0 / 255.0, 0.5, 0, 0)


def _empty_scene():
    z = np.zeros
    return GaussianSceneSoa(z((0, 3), np.float32), z((0, 6), np.float32), z(0, np.float32),
                            z((0, 48), np.float32))


def _states(scenes, cages=None):
    out = []
    for i, sc in enumerate(scenes):
        st = GpuPipelineState(0)
        st.sync_scene_and_cages(sc, None if cages is None else cages[i])
        out.append(st)
    return out


def _per_view(states, view_scene, cams, comps, p):
    imgs, counts = [], []
    for v in range(len(cams)):
        st = states[view_scene[v]]
        t = StageTimings()
        st.execute_frame_resident(cams[v], comps[v], p, t)
        W, H = max(1, comps[v].screen_width), max(1, comps[v].screen_height)
        imgs.append(st.readback_framebuffer_rgba8(W, H))
        counts.append((int(t.total_tile_entries), int(t.surviving_gaussians)))
    return imgs, counts


def _check_batch(batch, states, view_scene, cams, comps, p, label):
    stats = batch.execute(states, view_scene, cams, comps, p)
    print(f"[case] {label}: {stats.as_dict()}", flush=True)
    assert stats.views == len(cams)
    imgs, counts = _per_view(states, view_scene, cams, comps, p)
    for v in range(len(cams)):
        got = batch.readback_rgba8(v)
        assert got.shape == imgs[v].shape
        nd = int(np.count_nonzero(got != imgs[v]))
        assert nd == 0, f"{label}: view {v} differs from the per-view frame in {nd} bytes"
        if states[view_scene[v]]._keep and states[view_scene[v]]._keep[0].count:
            assert batch.view_counts(v) == counts[v], f"{label}: view {v} counts"
    assert stats.total_tile_entries == sum(c[0] for c in counts)
    return stats


def test_batch_matches_per_view_frames():
    scenes = [syn.scene_gaussian_ball(30_000, k=31), syn.scene_clusters(50_000, k=32),
              _empty_scene()]
    states = _states(scenes)
    view_scene, cams, comps = syn.batch_views(7, 3, seed=11, resolutions=SMALL)
    p = _params()
    with ViewBatch(0, lanes=3) as batch:
        s1 = _check_batch(batch, states, view_scene, cams, comps, p, "batch first")
        assert s1.graph_built == 1 and s1.lanes == 3
        # the second call may re-deal the views to lanes by their measured
        # weights (one more capture); from then on the graph is replayed
        _check_batch(batch, states, view_scene, cams, comps, p, "batch second")
        s2 = _check_batch(batch, states, view_scene, cams, comps, p, "batch replay")
        assert s2.graph_built == 0, "unchanged batch re-captured its graph"
        cams2 = list(cams)
        cams2[1] = syn.camera_look_at((0.3, 0.2, -3.5), (0, 0, 0), (0, 1, 0), 60.0,
                                      comps[1].screen_width, comps[1].screen_height)
        s3 = _check_batch(batch, states, view_scene, cams2, comps, p, "batch changed camera")
        assert s3.graph_built == 1
    for st in states:
        st.close()


def test_batch_lane_counts_agree():
    """1, 2 and 5 lanes give the same pixels (lanes only change concurrency)."""
    scenes = [syn.scene_gaussian_ball(20_000, k=33), syn.scene_gaussian_ball(25_000, k=34)]
    states = _states(scenes)
    view_scene, cams, comps = syn.batch_views(6, 2, seed=12, resolutions=SMALL)
    p = _params()
    ref = None
    for lanes in (1, 2, 5):
        with ViewBatch(0, lanes=lanes) as batch:
            st = batch.execute(states, view_scene, cams, comps, p)
            print(f"[case] lanes={lanes}: {st.as_dict()}", flush=True)
            imgs = [batch.readback_rgba8(v) for v in range(len(cams))]
        if ref is None:
            ref = imgs
        else:
            for v in range(len(cams)):
                assert np.array_equal(imgs[v], ref[v]), f"lanes={lanes} view {v}"
    for st in states:
        st.close()


def test_shared_projection_on_and_off_agree():
    """Stage 2 as one pass over each scene (default) or one pass per view: the
    same pixels and counts."""
    scenes = [syn.scene_gaussian_ball(20_000, k=38), syn.scene_clusters(30_000, k=39)]
    states = _states(scenes)
    view_scene, cams, comps = syn.batch_views(9, 2, seed=15, resolutions=SMALL)
    p = _params()
    out = {}
    for shared in (True, False):
        with ViewBatch(0, lanes=3) as batch:
            batch.set_shared_projection(shared)
            st = batch.execute(states, view_scene, cams, comps, p)
            print(f"[case] shared_projection={shared}: {st.as_dict()}", flush=True)
            out[shared] = ([batch.readback_rgba8(v) for v in range(len(cams))],
                           [batch.view_counts(v) for v in range(len(cams))])
    for v in range(len(cams)):
        assert np.array_equal(out[True][0][v], out[False][0][v]), f"view {v}"
        assert out[True][1][v] == out[False][1][v], f"view {v} counts"
    for st in states:
        st.close()


def test_batch_view_against_oracle(oracle_lib):
    import oracle

    scene = syn.scene_gaussian_ball(20_000, k=35)
    states = _states([scene])
    view_scene, cams, comps = syn.batch_views(3, 1, seed=13, resolutions=((320, 240),))
    with ViewBatch(0, lanes=2) as batch:
        batch.execute(states, view_scene, cams, comps, _params())
        for v in range(3):
            ref = oracle.render_frame(scene, cams[v], comps[v])
            got = batch.readback_rgba8(v)
            d = np.abs(got.astype(np.int16) - ref["rgba"].astype(np.int16))
            print(f"[case] oracle view {v}: E={ref['total']} max|d|={d.max()}", flush=True)
            assert batch.view_counts(v) == (ref["total"], ref["survivors"])
            assert d.max() <= 1
    for st in states:
        st.close()


def test_batch_deformed_scene_and_8k_view():
    from cindergs_amd.cages import CageSpringSimulator, TetMesh

    scene = syn.scene_gaussian_ball(30_000, k=36)
    lo, hi = scene.positions.min(0), scene.positions.max(0)
    mesh = TetMesh.jittered_grid(lo, hi, (8, 8, 8))
    tet, w = mesh.bind_gaussians(scene.positions)
    sim = CageSpringSimulator(mesh.rest_vertices, stiffness=40.0, damping=math.exp(-4.0 / 60.0))
    sim.impulse(0.2, 0.05)
    for _ in range(8):
        sim.step(1.0 / 60.0)
    cage = mesh.cage(sim.x, tet, w)
    plain = syn.scene_gaussian_ball(15_000, k=37)
    states = _states([scene, plain], cages=[cage, None])
    view_scene, cams, comps = syn.batch_views(4, 2, seed=14, resolutions=SMALL)
    # one 8K view (129,600 tiles: the 64-bit key path)
    cam8, comp8 = syn.camera_look_at((0.0, 0.0, -3.0), (0, 0, 0), (0, 1, 0), 60.0, 7680, 4320), \
        syn.compositor(7680, 4320, (0.0, 0.0, -3.0))
    view_scene.append(1)
    cams.append(cam8)
    comps.append(comp8)
    p = _params(tau=1.02)
    with ViewBatch(0, lanes=2) as batch:
        _check_batch(batch, states, view_scene, cams, comps, p, "deformed + 8K")
    for st in states:
        st.close()


def test_batch_presorts_views_with_many_entries_per_splat():
    """A view whose last replay had more than 4 tile entries per splat slot is
    sorted with the splats presorted by depth (cgs_execute_views_batched):
    the second call takes that path for the dense views, and every view stays
    bit-identical to its per-view frame; views with few entries per splat keep
    the LSD sort."""
    # a small dense ball seen large (many tiles per splat) beside a sparse view
    scenes = [syn.scene_gaussian_ball(6_000, k=37), syn.scene_clusters(40_000, k=38)]
    states = _states(scenes)
    view_scene = [0, 1, 0]
    cams, comps = [], []
    for eye_z, (w, h) in ((-1.2, (1920, 1080)), (-3.0, (320, 240)), (-2.0, (800, 600))):
        eye = (0.1, 0.05, eye_z)
        cams.append(syn.camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, w, h))
        comps.append(syn.compositor(w, h, eye))
    p = _params()
    with ViewBatch(0, lanes=2) as batch:
        s1 = _check_batch(batch, states, view_scene, cams, comps, p, "presort first")
        assert s1.presort_views == 0  # nothing measured yet
        s2 = _check_batch(batch, states, view_scene, cams, comps, p, "presort second")
        per_splat = [batch.view_counts(v)[0] / scenes[view_scene[v]].count for v in range(3)]
        print(f"[case] entries per splat {per_splat}", flush=True)
        assert s2.presort_views == sum(1 for r in per_splat if r > 4.0) >= 1
        s3 = _check_batch(batch, states, view_scene, cams, comps, p, "presort replay")
        assert s3.graph_built == 0 and s3.presort_views == s2.presort_views
    for st in states:
        st.close()
