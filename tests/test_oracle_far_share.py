"""The oracle's far share of MVC's sums (orc_mvc_far_share) and its far sums
(orc_mvc_far_sums, the same sums before the normalisation), pinned by what
the split fixes: with no triangle near the point (r1 below every centroid
distance) the far share is the exact weights (times nv), with every triangle
near it is zero, a larger near radius never adds far weight, the far share
sums to at most 1 (the cage is star-shaped about its centre, so no triangle's
terms are negative at interior points), and it is continuous in x; the far
sums normalise to the exact weights when no triangle is near, equal the
shares times the whole sum, and the whole sum is at least 4 pi / diagonal."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))

from cindergs_amd.mvc import icosphere


@pytest.fixture(scope="module")
def setup(oracle_lib):
    import oracle

    cage = icosphere(3, 1.2, 0.15)
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.7, 0.7, (40, 3)).astype(np.float32)
    return oracle, cage, x


def test_no_near_triangle_gives_the_exact_weights(setup):
    o, cage, x = setup
    q = o.mvc_far_share(cage.vertices, cage.triangles, x, 1e-6, 2e-6)
    w = o.mvc_weights(cage.vertices, cage.triangles, x, fp64=True)
    np.testing.assert_allclose(q, cage.num_vertices * w, atol=1e-12)


def test_all_triangles_near_gives_zero(setup):
    o, cage, x = setup
    q = o.mvc_far_share(cage.vertices, cage.triangles, x, 10.0, 20.0)
    assert not np.any(q)


def test_far_share_shrinks_with_the_near_radius(setup):
    o, cage, x = setup
    nv = cage.num_vertices
    sums = [o.mvc_far_share(cage.vertices, cage.triangles, x, r0, r1).sum(1) / nv
            for r0, r1 in ((0.1, 0.2), (0.2, 0.4), (0.3, 0.55), (0.5, 0.9))]
    for a, b in zip(sums, sums[1:]):
        assert np.all(b <= a + 1e-12)
    assert np.all(sums[0] <= 1.0 + 1e-12) and np.all(sums[-1] >= 0.0)


def test_far_share_is_continuous(setup):
    o, cage, x = setup
    d = np.float32(1e-3)
    q0 = o.mvc_far_share(cage.vertices, cage.triangles, x, 0.3, 0.55)
    q1 = o.mvc_far_share(cage.vertices, cage.triangles, x + d, 0.3, 0.55)
    # |dq/dx| is bounded by the terms' own derivatives: a 1e-3 step moves
    # each entry (O(1) after the nv scaling) by far less than 0.1
    assert np.abs(q1 - q0).max() < 0.1


# The far sums (orc_mvc_far_sums): the same sums before the normalisation, so
# they are pinned by the shares above and by the exact weights.


def test_far_sums_with_no_near_triangle_normalise_to_the_exact_weights(setup):
    o, cage, x = setup
    s = o.far_sums_scale(cage.vertices)
    S = o.mvc_far_sums(cage.vertices, cage.triangles, x, 1e-6, 2e-6, s)
    w = o.mvc_weights(cage.vertices, cage.triangles, x, fp64=True)
    np.testing.assert_allclose(S / S.sum(1, keepdims=True), w, atol=1e-12)
    # every interior point's MVC sum is at least 4 pi / diagonal, so the
    # scaled sums add up to at least nv
    assert np.all(S.sum(1) >= cage.num_vertices)


def test_far_sums_are_the_shares_times_the_whole_sum(setup):
    o, cage, x = setup
    nv, s = cage.num_vertices, o.far_sums_scale(cage.vertices)
    full = o.mvc_far_sums(cage.vertices, cage.triangles, x, 1e-6, 2e-6, s).sum(1, keepdims=True)
    for r0, r1 in ((0.2, 0.4), (0.3, 0.55)):
        S = o.mvc_far_sums(cage.vertices, cage.triangles, x, r0, r1, s)
        q = o.mvc_far_share(cage.vertices, cage.triangles, x, r0, r1)
        np.testing.assert_allclose(S, q * full / nv, rtol=1e-10, atol=1e-12)
    assert not np.any(o.mvc_far_sums(cage.vertices, cage.triangles, x, 10.0, 20.0, s))


def test_far_sums_scale_is_nv_diagonal_over_4pi(setup):
    o, cage, _ = setup
    assert o.far_sums_scale(cage.vertices) == pytest.approx(
        cage.num_vertices * cage.diagonal() / (4 * np.pi), rel=1e-6)
