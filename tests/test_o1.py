"""O1, the restatement of the reference's Rayon CPU render path
(oracle/cgs_o1.cpp), against the reference's own known answers.

* The golden image: render_scene_cpu of fibonacci_cluster(256, 0.5, [0,0,2])
  with the identity camera (f = 180, c = (80, 60), 160x120, near 0.2),
  background [0.05, 0.05, 0.07, 1], SH degree 2 hashes (FNV-1a over RGBA8) to
  0xbdd6_3815_3115_0f06 (tests/gpu_parity_tests.rs:293-312), on 1 thread and on
  several, and re-rendering gives the same image.
* Projection known answers (src/math/projection.rs:556-740).
* Binning fixtures (src/render/tiles/cpu_reference.rs:408-535) and a
  splat-by-splat enumeration sorted by Python, which O1's binary-search
  scatter and parallel sort must equal.
* Rasterizer fixtures (src/render/rasterizer/cpu_reference.rs:247-358).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))

GOLDEN = 0xBDD6_3815_3115_0F06


@pytest.fixture(scope="module")
def o(oracle_lib):
    import oracle

    return oracle


def _eye():
    return np.eye(4, dtype=np.float32).reshape(16)


def _golden(o, threads):
    pos, cov, opa, sh = o.fibonacci_cluster(256, 0.5, [0.0, 0.0, 2.0])
    cam = o.o1_camera(_eye(), 180.0, 180.0, 80.0, 60.0, 160.0, 120.0, 0.2)
    img, stats = o.o1_render(pos, cov, opa, sh, cam, 160, 120, [0.05, 0.05, 0.07, 1.0],
                             [0.0, 0.0, 0.0], 2, threads=threads)
    return img, stats


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_golden_image_hash(o, threads):
    img, stats = _golden(o, threads)
    assert stats["surviving"] == 256
    assert o.pixel_hash(img) == GOLDEN, hex(o.pixel_hash(img))
    again, _ = _golden(o, threads)
    assert np.array_equal(img, again)


def test_golden_hash_detects_a_changed_pixel(o):
    img, _ = _golden(o, 2)
    img = img.copy()
    img[60, 80, 0] ^= 1
    assert o.pixel_hash(img) != GOLDEN


def test_fibonacci_cluster_shape(o):
    # synthetic.rs:47-60: every splat within the radius of the centre
    pos, cov, opa, sh = o.fibonacci_cluster(128, 1.0, [0.0, 1.0, 2.0])
    d = np.linalg.norm(pos.astype(np.float64) - [0.0, 1.0, 2.0], axis=1)
    assert d.max() <= 1.0 + 1e-4
    assert np.all(opa == np.float32(0.88))
    assert np.all(cov == np.float32([0.0004, 0, 0, 0.0004, 0, 0.0004]))
    assert np.all(sh[:, 3:] == 0.0)


def _project_one(o, pos, cov, f, c, wh, near=0.2):
    cam = o.o1_camera(_eye(), f[0], f[1], c[0], c[1], wh[0], wh[1], near)
    pts, d, r, con = o.o1_project(np.float32([pos]), np.float32([cov]), cam, threads=1)
    return pts[0], d[0], r[0], con[0]


def test_projection_manual_reference(o):
    # projection.rs:556-597
    pts, d, r, con = _project_one(o, [0, 0, 2], [0.04, 0, 0, 0.04, 0, 0.04], (1000, 1000),
                                  (500, 400), (1000, 800))
    assert d == 2.0 and list(pts) == [500.0, 400.0]
    assert r == 301
    assert abs(con[0] - 1.0 / 10000.3) < 1e-6 and con[1] == 0.0
    assert abs(con[2] - 1.0 / 10000.3) < 1e-6


@pytest.mark.parametrize("z", [0.1, 0.2, -5.0])
def test_near_plane_culling(o, z):
    # projection.rs:600-654
    _, _, r, _ = _project_one(o, [0, 0, z], [0.01, 0, 0, 0.01, 0, 0.01], (500, 500), (250, 250),
                              (500, 500))
    assert r == 0


def test_frustum_culling_outside_viewport(o):
    # projection.rs:657-677
    _, _, r, _ = _project_one(o, [100, 0, 2], [0.01, 0, 0, 0.01, 0, 0.01], (500, 500), (250, 250),
                              (500, 500))
    assert r == 0


@pytest.mark.parametrize("cov", [[0.1, 0.02, 0.01, 0.08, 0.03, 0.15], [1, 0, 0, 1, 0, 1],
                                 [0.001, 0, 0, 0.001, 0, 0.001]])
def test_conic_positive_definite(o, cov):
    # projection.rs:680-718
    _, _, r, (a, b, c) = _project_one(o, [0, 0, 3], cov, (800, 800), (400, 300), (800, 600))
    assert r > 0 and a > 0 and c > 0 and a * c - b * b > 0


def test_radius_clamped_at_1024(o):
    # projection.rs:721-740
    _, _, r, _ = _project_one(o, [0, 0, 1], [1000, 0, 0, 1000, 0, 1000], (1000, 1000), (500, 500),
                              (2000, 2000))
    assert r == 1024


def _bin(o, pts, depths, radii, W=800, H=600):
    return o.o1_bin(np.float32(pts).reshape(-1, 2), np.float32(depths), np.int32(radii), W, H,
                    threads=4)


def test_binning_fixtures(o):
    # cpu_reference.rs:408-493
    e, rng = _bin(o, [8, 8], [2.5], [2])
    assert e.tolist() == [[0, np.float32(2.5).view(np.uint32), 0]]
    assert rng[0].tolist() == [0, 1] and rng[1].tolist() == [0, 0]
    e, rng = _bin(o, [16, 16], [3.0], [2])
    assert e[:, 0].tolist() == [0, 1, 50, 51] and e[:, 2].tolist() == [0] * 4
    assert int(np.sum(rng[:, 1] > rng[:, 0])) == 4
    e, rng = _bin(o, [100, 100], [1.0], [0])
    assert e.shape[0] == 0 and not np.any(rng)
    e, rng = _bin(o, [4, 4, 8, 8, 24, 8], [5.0, 1.0, 2.0], [2, 2, 2])
    assert e[:, 0].tolist() == [0, 0, 1] and e[:, 2].tolist() == [1, 0, 2]
    assert e[:, 1].view(np.float32).tolist() == [1.0, 5.0, 2.0]
    assert rng[0].tolist() == [0, 2] and rng[1].tolist() == [2, 3]


def _enumerate(pts, depths, radii, W, H):
    """Splat-major enumeration of each splat's tile rectangle, sorted by Python."""
    tx, ty = (W + 15) // 16, (H + 15) // 16
    out = []
    for g in range(len(radii)):
        r = float(radii[g])
        if r <= 0:
            continue
        cx, cy = float(pts[g, 0]), float(pts[g, 1])
        if cx + r < 0 or cx - r >= W or cy + r < 0 or cy - r >= H:
            continue
        x0 = min(max(int(np.floor(np.float32(cx - r) / 16)), 0), tx - 1)
        x1 = min(max(int(np.floor(np.float32(cx + r) / 16)), 0), tx - 1)
        y0 = min(max(int(np.floor(np.float32(cy - r) / 16)), 0), ty - 1)
        y1 = min(max(int(np.floor(np.float32(cy + r) / 16)), 0), ty - 1)
        d = int(np.float32(depths[g]).view(np.uint32))
        out += [(y * tx + x, d, g) for y in range(y0, y1 + 1) for x in range(x0, x1 + 1)]
    return sorted(out)


def test_scatter_and_sort_equal_direct_enumeration(o):
    rng = np.random.default_rng(7)
    n, W, H = 3000, 333, 217
    pts = rng.uniform(-40, [W + 40, H + 40], size=(n, 2)).astype(np.float32)
    depths = rng.choice(rng.uniform(0.3, 9.0, 200), n).astype(np.float32)  # repeated depths
    radii = rng.integers(-2, 40, n).astype(np.int32)
    e, ranges = o.o1_bin(pts, depths, radii, W, H, threads=8)
    want = _enumerate(pts, depths, radii, W, H)
    assert [tuple(x) for x in e.tolist()] == want
    T = ranges.shape[0]
    counts = np.bincount(e[:, 0], minlength=T)
    nz = counts > 0
    assert np.array_equal((ranges[:, 1] - ranges[:, 0])[nz], counts[nz])
    assert not np.any(ranges[~nz])
    e1, r1 = o.o1_bin(pts, depths, radii, W, H, threads=1)
    assert np.array_equal(e, e1) and np.array_equal(ranges, r1)


def _raster(o, pts, conics, opa, sh, pos, entries, ranges, W=16, H=16, bg=(0, 0, 0, 1), deg=0):
    return o.o1_rasterize(np.float32(pts), np.float32(conics), np.float32(opa), np.float32(sh),
                          np.float32(pos), np.uint32(entries).reshape(-1, 3),
                          np.uint32(ranges).reshape(-1, 2), W, H, bg, [0, 0, 0], deg, threads=2)


def test_rasterizer_fixtures(o):
    f2 = int(np.float32(2.0).view(np.uint32))
    sh = np.zeros(48, np.float32)
    sh[0] = 1.0
    _, f = _raster(o, [0.5, 0.5], [1, 0, 1], [1.0], sh, [0, 0, 2], [0, f2, 0], [0, 1])
    assert f[0, 0, 0] > 0.7  # cpu_reference.rs:248-281
    sh = np.zeros((5, 48), np.float32)
    sh[0, :3] = (2.0, -2.0, -2.0)
    sh[1:, :3] = (-2.0, 2.0, -2.0)
    entries = [[0, int(np.float32(i).view(np.uint32)), i] for i in range(5)]
    _, f = _raster(o, [0.5, 0.5] * 5, [1, 0, 1] * 5, [0.99] * 5, sh, np.zeros(15), entries, [0, 5])
    assert f[0, 0, 0] > 0.9 and f[0, 0, 1] < 0.1  # cpu_reference.rs:284-331
    img, f = _raster(o, [], [], [], np.zeros(48), [], [0, 0, 0], [0, 0] * 4, 19, 19, (0.1, 0.2, 0.3, 1))
    assert img.shape == (19, 19, 4)
    assert abs(f[18, 18, 0] - 0.1) < 1e-5 and abs(f[18, 18, 1] - 0.2) < 1e-5  # 334-348
    a, _ = _raster(o, [], [], [], np.zeros(48), [], [0, 0, 0], [0, 0], 8, 8, (0.1, 0.2, 0.3, 1))
    assert o.pixel_hash(a) != 0  # 351-357


def _real_sh_basis(d):
    """Real SH of degree <= 3 built from scipy's complex SH (Condon-Shortley
    phase): m < 0 -> sqrt2 Im Y_l^|m|, m = 0 -> Y_l^0, m > 0 -> sqrt2 Re Y_l^m,
    ordered m = -l..l.  This is the basis sh.rs:5-137 writes out by hand."""
    from scipy.special import sph_harm_y

    x, y, z = d
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    out = []
    for l in range(4):
        for m in range(-l, l + 1):
            Y = sph_harm_y(l, abs(m), theta, phi)
            out.append(np.sqrt(2.0) * Y.imag if m < 0 else (Y.real if m == 0 else np.sqrt(2.0) * Y.real))
    return np.array(out)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_evaluate_sh_equals_the_real_sh_basis(o, degree):
    rng = np.random.default_rng(11 + degree)
    nb = (degree + 1) ** 2
    for _ in range(40):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        c = np.zeros(48, np.float32)
        c[:3] = rng.uniform(-0.4, 0.4, 3)
        rest = rng.uniform(-0.15, 0.15, (15, 3)).astype(np.float32)
        for ch in range(3):
            c[3 + 15 * ch:18 + 15 * ch] = rest[:, ch]
        coeff = np.vstack([c[:3][None, :], rest])  # [16, 3] per basis, per channel
        B = _real_sh_basis(d)
        want = np.clip(0.5 + B[:nb] @ coeff[:nb].astype(np.float64), 0.0, 1.0)
        got = o.o1_evaluate_sh(c, d.astype(np.float32), degree)
        np.testing.assert_allclose(got, want, atol=2e-6)


def test_evaluate_sh_reference_cases(o):
    # sh.rs:143-181
    sh = np.zeros(48, np.float32)
    assert o.o1_evaluate_sh(sh, [0, 0, 1], 0).tolist() == [0.5, 0.5, 0.5]
    sh[0] = 1.0
    r = o.o1_evaluate_sh(sh, [0, 0, 1], 0)
    assert abs(r[0] - (0.5 + 0.28209479177387814)) < 1e-6 and r[1] == 0.5 and r[2] == 0.5
    sh[0], sh[1] = 100.0, -100.0
    r = o.o1_evaluate_sh(sh, [0, 0, 1], 0)
    assert r[0] == 1.0 and r[1] == 0.0
    sh[:] = 0.0
    sh[4] = 1.0
    assert abs(o.o1_evaluate_sh(sh, [0, 0, 1], 1)[0] - (0.5 + 0.4886025119029199)) < 1e-6
    assert abs(o.o1_evaluate_sh(sh, [0, 0, -1], 1)[0] - (0.5 - 0.4886025119029199)) < 1e-6


def _dc(rgb):
    """DC coefficients giving colour `rgb` under evaluate_sh."""
    return [(v - 0.5) / 0.28209479177387814 for v in rgb]


def test_rasterizer_termination_clamp_and_cutoffs(o):
    # Front-to-back: stop before a splat once T < 1e-3 and set T = 0
    # (cpu_reference.rs:171-175); alpha = min(o exp(q), 0.99) (196); skip
    # q < -4 (190-192) and alpha < 1/255 (197-199).
    cols = [(0.8, 0.2, 0.1), (0.3, 0.6, 0.9), (0.5, 0.5, 0.5)]
    sh = np.zeros((3, 48), np.float32)
    for i, c in enumerate(cols):
        sh[i, :3] = _dc(c)
    entries = [[0, int(np.float32(i + 1).view(np.uint32)), i] for i in range(3)]
    bg = (0.1, 0.2, 0.3, 1.0)
    _, f = _raster(o, [0.5, 0.5] * 3, [1, 0, 1] * 3, [1.0] * 3, sh, np.zeros(9), entries, [0, 3],
                   bg=bg)
    # alpha 0.99, 0.99 -> T = 1e-4 < 1e-3 before the third splat: T = 0
    want = [0.99 * cols[0][k] + 0.99 * 0.01 * cols[1][k] for k in range(3)] + [1.0]
    np.testing.assert_allclose(f[0, 0], want, atol=2e-6)
    # one splat: alpha clamped at 0.99, background behind it
    _, f = _raster(o, [0.5, 0.5], [1, 0, 1], [1.0], sh[:1], np.zeros(3), entries[:1], [0, 1], bg=bg)
    np.testing.assert_allclose(f[0, 0], [0.99 * cols[0][k] + 0.01 * bg[k] for k in range(3)] + [0.99],
                               atol=2e-6)
    # q just inside / outside the -4 cut: dx^2 / 2 = 3.92 / 4.205 (a = c = 1)
    for dx, inside in ((2.8, True), (2.9, False)):
        _, f = _raster(o, [0.5 + dx, 0.5], [1, 0, 1], [1.0], sh[:1], np.zeros(3), entries[:1], [0, 1],
                       bg=bg)
        a = np.float32(np.exp(np.float32(-0.5 * dx * dx))) if inside else 0.0
        assert abs(f[0, 0, 3] - a) < 1e-6, (dx, f[0, 0, 3])
    # opacity just above / below 1/255 at q = 0
    for op, inside in ((0.004, True), (0.0039, False)):
        _, f = _raster(o, [0.5, 0.5], [1, 0, 1], [op], sh[:1], np.zeros(3), entries[:1], [0, 1], bg=bg)
        assert abs(f[0, 0, 3] - (op if inside else 0.0)) < 1e-7, (op, f[0, 0, 3])
