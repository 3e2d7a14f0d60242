"""The synthetic workloads' generator (cindergs_amd.rng and its users).

Pins:
* SplitMix64 and xoshiro256** against their published outputs (splitmix64.c
  from seed 0; xoshiro256** from the state {1, 2, 3, 4});
* the vectorised streams against a scalar Python-integer restatement of the
  spec: the stream seeding of (seed, domain, index), uniforms and Box-Muller
  normals, and the draw order when uniforms and normals are mixed;
* chunked and threaded draws against one draw of all streams;
* the first records of every workload (configs 1-5) against the committed
  fixtures of tests/golden/make_generator_fixtures.py, bit for bit;
* the moments of the normal and uniform draws.
"""
import json
import math
import os

import numpy as np
import pytest

from cindergs_amd import rng as R
from cindergs_amd import synthetic as syn

HERE = os.path.dirname(os.path.abspath(__file__))
M = (1 << 64) - 1


def _sm(z):
    z = (z + 0x9E3779B97F4A7C15) & M
    x = z
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
    return z, x ^ (x >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M


class _Scalar:
    """One stream of the spec, in Python integers."""

    def __init__(self, seed, domain, index):
        z = (seed + 0x9E3779B97F4A7C15 * 4 * ((domain << 32) + index)) & M
        self.s = []
        for _ in range(4):
            z, w = _sm(z)
            self.s.append(w)

    def next(self):
        s0, s1, s2, s3 = self.s
        out = (_rotl((s1 * 5) & M, 7) * 9) & M
        t = (s1 << 17) & M
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = _rotl(s3, 45)
        self.s = [s0, s1, s2, s3]
        return out

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def normal_pair(self):
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(1.0 - u1))
        return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)


def test_published_outputs():
    z, outs = 0, []
    for _ in range(4):
        z, o = _sm(z)
        outs.append(o)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                    0xF88BB8A8724C81EC]
    st = R.Streams(0, 0, 1, state=[1, 2, 3, 4])
    assert [int(st.next_u64()[0]) for _ in range(4)] == [11520, 0, 1509978240,
                                                        1215971899390074240]
    z = np.zeros(1, np.uint64)
    for want in outs:
        z, o = R.splitmix64_next(z)
        assert int(o[0]) == want


@pytest.mark.parametrize("seed,domain,first", [(0xC1DE5EED00000002, 1, 0),
                                               (0xC1DE5EED00000005, 2, 61),
                                               (M, 6, (1 << 32) - 3)])
def test_streams_match_the_scalar_spec(seed, domain, first):
    n = 5
    st = R.Streams(seed, domain, n, first)
    u = st.uniform(3)
    z = st.normal(5)
    for j in range(n):
        sc = _Scalar(seed, domain, first + j)
        for k in range(3):
            assert u[j, k] == sc.uniform()
        got = []
        for _ in range(3):
            got += list(sc.normal_pair())
        np.testing.assert_allclose(z[j], got[:5], rtol=0, atol=1e-15)


def test_chunked_draws_equal_one_draw():
    seed = 0xC1DE5EED00000009
    whole = R.Streams(seed, 1, 200_000).normal(7)
    parts = R.chunked(seed, 1, 200_000, lambda st: st.normal(7), chunk=30_000)
    assert np.array_equal(whole, parts)
    off = R.Streams(seed, 1, 100, first=123_456).normal(7)
    assert np.array_equal(off, whole[123_456:123_556])


def test_moments():
    st = R.Streams(0xC1DE5EED00000003, 1, 400_000)
    z = st.normal(2)
    u = st.uniform(1)
    assert abs(z.mean()) < 5e-3 and abs(z.var() - 1.0) < 1e-2
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 2e-3


@pytest.fixture(scope="module")
def fix():
    with open(os.path.join(HERE, "golden", "generator_fixtures.json")) as f:
        return json.load(f)


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).ravel().tolist()


def _check_scene(scene, rec):
    lo, hi = rec["first"], rec["first"] + rec["count"]
    assert _bits(scene.positions[lo:hi]) == rec["positions"]
    assert _bits(scene.covariances_3d[lo:hi]) == rec["covariances"]
    assert _bits(scene.opacities[lo:hi]) == rec["opacities"]
    assert _bits(scene.sh_coeffs[lo:hi]) == rec["sh"]


def test_config1_first_splats(fix):
    _check_scene(syn.scene_gaussian_ball(), fix["config1_ball_100k"])


def test_config2_first_members_and_background(fix):
    scene = syn.scene_clusters()
    _check_scene(scene, fix["config2_clusters_members"])
    _check_scene(scene, fix["config2_clusters_background"])


def test_config3_grid(fix):
    from cindergs_amd.cages import TetMesh

    rec = fix["config3_grid_vertices"]
    mesh = TetMesh.jittered_grid(rec["lo"], rec["hi"], tuple(rec["cells"]), rec["jitter"],
                                 rec["seed"])
    assert _bits(mesh.rest_vertices[:rec["count"]]) == rec["vertices"]


def test_config4_icosphere(fix):
    from cindergs_amd.mvc import icosphere

    rec = fix["config4_icosphere_vertices"]
    assert _bits(icosphere().vertices[:rec["count"]]) == rec["vertices"]


def test_config5_views(fix):
    rec = fix["config5_views"]
    vs, cams, comps = syn.batch_views(rec["count"], 3, seed=5)
    assert [int(v) for v in vs] == rec["scene"]
    assert [[c.screen_width, c.screen_height] for c in comps] == rec["size"]
    assert _bits([list(c.camera_position) for c in comps]) == rec["eye"]
