"""SortCache (src/render/pipeline.rs:40-63, 66-193) through the C ABI, on the
CPU: the reference's three unit tests (pipeline.rs:1071-1113) and the edges
they leave open (the 1e-5 bound itself, the frame-parameter bits, cage
length changes, a miss dropping the binning, the re-anchoring commit)."""
import numpy as np
import pytest

from cindergs_amd.pipeline import SortCache

EPS = np.float32(1e-5)


@pytest.fixture
def cache():
    c = SortCache()
    yield c
    c.close()


def test_sort_cache_reuses_when_nothing_moves(cache):
    """pipeline.rs:1071-1081."""
    view = np.zeros(16, np.float32)
    cage = [[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]]
    assert not cache.can_reuse(view, cage, 800, 600)
    cache.commit(view, cage, 800, 600, "binning")
    assert cache.can_reuse(view, cage, 800, 600)
    assert cache.take_if_reusable(view, cage, 800, 600) == "binning"
    assert cache.reuse_count == 1


def test_sort_cache_invalidates_on_camera_or_cage_motion(cache):
    """pipeline.rs:1084-1100."""
    view = np.zeros(16, np.float32)
    cage = [[0.0, 0.0, 0.0]]
    cache.commit(view, cage, 800, 600, "binning")
    moved = view.copy()
    moved[12] = 0.5
    assert not cache.can_reuse(moved, cage, 800, 600)
    assert not cache.can_reuse(view, [[0.5, 0.0, 0.0]], 800, 600)
    assert not cache.can_reuse(view, cage, 1024, 768)


def test_sort_cache_tiny_motion_is_below_threshold(cache):
    """pipeline.rs:1103-1113."""
    view = np.zeros(16, np.float32)
    cache.commit(view, [[0.0, 0.0, 0.0]], 800, 600, "binning")
    jittered = view.copy()
    jittered[0] = EPS * np.float32(0.5)
    assert cache.can_reuse(jittered, [[EPS * np.float32(0.5), 0.0, 0.0]], 800, 600)


def test_the_bound_is_inclusive_and_in_f32(cache):
    view = np.zeros(16, np.float32)
    cache.commit(view, [[0.0, 0.0, 0.0]], 800, 600, "b")
    at = view.copy()
    at[5] = EPS
    assert cache.can_reuse(at, [[0.0, 0.0, 0.0]], 800, 600)
    above = view.copy()
    above[5] = np.nextafter(EPS, np.float32(1.0))
    assert not cache.can_reuse(above, [[0.0, 0.0, 0.0]], 800, 600)
    assert cache.can_reuse(view, [[0.0, -EPS, 0.0]], 800, 600)
    assert not cache.can_reuse(view, [[0.0, 0.0, -np.nextafter(EPS, np.float32(1.0))]], 800, 600)


def test_cage_count_change_never_reuses(cache):
    view = np.zeros(16, np.float32)
    cache.commit(view, [[0.0, 0.0, 0.0]], 64, 64, "b")
    assert not cache.can_reuse(view, [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]], 64, 64)
    assert not cache.can_reuse(view, np.zeros((0, 3)), 64, 64)
    cache.commit(view, np.zeros((0, 3)), 64, 64, "b")
    assert cache.can_reuse(view, np.zeros((0, 3)), 64, 64)


def test_for_frame_keys_on_the_parameter_bits(cache):
    """take_if_reusable_for_frame (pipeline.rs:119-145): splat count, focal
    lengths, split threshold and mip variance compared bit for bit (so 0.0 and
    -0.0 differ)."""
    view = np.eye(4, dtype=np.float32).reshape(16)
    cage = np.zeros((4, 3), np.float32)
    base = dict(splat_count=1000, focal_x=519.6, focal_y=519.6, split_threshold=0.0,
                mip_var=0.0)
    cache.commit_for_frame(view, cage, 800, 600, binning="b", **base)
    assert cache.take_if_reusable_for_frame(view, cage, 800, 600, **base) == "b"
    for key, value in (("splat_count", 1001), ("focal_x", np.nextafter(np.float32(519.6), 1e9)),
                       ("focal_y", 519.0), ("split_threshold", -0.0), ("mip_var", 1e-4)):
        cache.commit_for_frame(view, cage, 800, 600, binning="b", **base)
        other = dict(base)
        other[key] = value
        assert cache.take_if_reusable_for_frame(view, cage, 800, 600, **other) is None, key
        # a miss drops the binning: nothing to take afterwards, even for a match
        assert cache.take_if_reusable_for_frame(view, cage, 800, 600, **base) is None
    assert cache.reuse_count == 1


def test_take_removes_the_binning_and_commit_reanchors(cache):
    """A reuse takes the binning out; the caller commits it back with the new
    view, so motion below 1e-5 per frame keeps reusing (execute_cpu_fallback_frame,
    pipeline.rs:552-600)."""
    view = np.zeros(16, np.float32)
    cage = [[0.0, 0.0, 0.0]]
    cache.commit(view, cage, 32, 32, "b")
    for step in range(1, 6):
        v = view.copy()
        v[3] = np.float32(step) * np.float32(0.6e-5)
        b = cache.take_if_reusable(v, cage, 32, 32)
        assert b == "b"
        assert not cache.can_reuse(v, cage, 32, 32)  # taken
        cache.commit(v, cage, 32, 32, b)
    assert cache.reuse_count == 5
    far = view.copy()
    far[3] = np.float32(5) * np.float32(0.6e-5) + np.float32(2e-5)
    assert cache.take_if_reusable(far, cage, 32, 32) is None
    cache.invalidate()
    assert not cache.can_reuse(view, cage, 32, 32)


def test_nan_motion_is_ignored_like_f32_max(cache):
    """f32::max returns the other operand for NaN, so a NaN difference does not
    block reuse (the reference's max_abs_difference)."""
    view = np.zeros(16, np.float32)
    cache.commit(view, [[0.0, 0.0, 0.0]], 8, 8, "b")
    v = view.copy()
    v[7] = np.nan
    assert cache.can_reuse(v, [[0.0, 0.0, 0.0]], 8, 8)
