"""The oracle pinned against the reference's own known answers (CPU only).

Every case comes from tests/golden/reference_fixtures.json, which holds the
inputs and expected outputs of the reference's unit tests (file:line in each
entry), or from what the mathematics fixes: closed forms of the EWA
projection, libm to within a few ulp for the shared elementary functions,
numpy's stable sort and cumulative sum for the scan and radix sort, a
brute-force pixel enumeration for the ellipse-vs-tile test, and polar
decomposition identities.  The oracle's FP64 twin is checked alongside.
"""
import json
import math
import os

import numpy as np
import pytest

from cindergs_amd.pipeline import GaussianSceneSoa, TetCage

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))
F32 = np.float32


def _camera24(view_rowmajor, fx, fy, cx, cy, W, H, near):
    """CameraUniforms (camera.rs:9-36): view[16] column-major, fx, fy, cx, cy,
    W, H, near, pad."""
    v = np.asarray(view_rowmajor, dtype=np.float32).reshape(4, 4)
    return np.concatenate([v.T.reshape(-1), np.array([fx, fy, cx, cy, W, H, near, 0.0], np.float32)])


def _scene(pos, cov, opa, sh=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    return GaussianSceneSoa(pos, np.asarray(cov, np.float32).reshape(n, 6),
                            np.asarray(opa, np.float32).reshape(n),
                            np.zeros((n, 48), np.float32) if sh is None else sh)


def _ulp_err(got32, exact64):
    got = np.asarray(got32, np.float64)
    spacing = np.spacing(np.abs(exact64).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(got - exact64) / spacing)


# ---------------------------------------------------------------------------
# elementary functions shared by the oracle and the kernels (cgs_detmath.h)
# ---------------------------------------------------------------------------

def test_detmath_matches_libm_to_a_few_ulp(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(1e-6, 1e3, 4000), np.geomspace(1e-30, 1e30, 2000),
                         [1.0, 2.0, 255.0, 0.5]]).astype(np.float32)
    got = np.array([L.orc_logf(float(x)) for x in xs], np.float32)
    assert _ulp_err(got, np.log(xs.astype(np.float64))) <= 2.0
    xs = rng.uniform(-87.0, 88.0, 6000).astype(np.float32)
    got = np.array([L.orc_expf(float(x)) for x in xs], np.float32)
    assert _ulp_err(got, np.exp(xs.astype(np.float64))) <= 2.0
    xs = rng.uniform(1e-4, 1.0, 3000).astype(np.float32)
    ys = rng.uniform(0.2, 3.0, 3000).astype(np.float32)
    got = np.array([L.orc_powf(float(x), float(y)) for x, y in zip(xs, ys)], np.float32)
    # pow = exp(y ln x): the ln error is scaled by y |ln x| (<= 3 * 9.2 here)
    assert _ulp_err(got, np.power(xs.astype(np.float64), ys.astype(np.float64))) <= 64.0
    xs = np.concatenate([rng.uniform(-1.0, 1.0, 4000), [-1.0, 0.0, 1.0]]).astype(np.float32)
    got = np.array([L.orc_acosf(float(x)) for x in xs], np.float32)
    err = np.abs(got.astype(np.float64) - np.arccos(xs.astype(np.float64)))
    assert err.max() <= 4e-7 * math.pi
    xs = rng.uniform(-math.pi, math.pi, 4000).astype(np.float32)
    got = np.array([L.orc_cosf(float(x)) for x in xs], np.float32)
    # Taylor series to x^18 evaluated in f32: the alternating terms near |x| = pi
    # cancel, about 7 ulp of |cos x| at worst
    assert np.abs(got.astype(np.float64) - np.cos(xs.astype(np.float64))).max() <= 5e-7


def test_half_to_float_is_exact_for_every_half(oracle_lib):
    L = oracle_lib.lib()
    h = np.arange(0, 1 << 16, 7, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([L.orc_half_to_float(int(v)) for v in h], np.float32)
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(got[fin].view(np.uint32), ref[fin].view(np.uint32))
    assert np.all(np.isnan(got[np.isnan(ref)]))


# ---------------------------------------------------------------------------
# Stage 2: projection closed forms (projection.rs:555-745)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", FIX["projection"], ids=lambda c: c["name"])
@pytest.mark.parametrize("fp64", [False, True])
def test_projection_known_answers(oracle_lib, case, fp64):
    cam = _camera24(np.eye(4), case["fx"], case["fy"], case["cx"], case["cy"], case["W"],
                    case["H"], case["near"])
    sc = _scene(case["pos"], case["cov"], case["opacity"])
    attrs, ext, surv = oracle_lib.project(sc, cam, [0.0, 0.0, 0.0], sh_degree=0, fp64=fp64)
    e = case["expect"]
    if e["culled"]:
        assert surv == 0 and ext[0] == 0
        return
    assert surv == 1 and ext[0] != 0
    a = attrs[0]
    ex, ey = int(ext[0]) & 0xFFFF, int(ext[0]) >> 16
    if "mu" in e:
        assert a[0] == e["mu"][0] and a[1] == e["mu"][1]
        assert a[9] == e["depth"]
        assert abs(a[2] - e["conic_a"]) < e["conic_tol"]
        assert a[3] == 0.0
        assert abs(a[4] - e["conic_c"]) < e["conic_tol"]
        # closed form of the EWA footprint: sigma^2 = f^2 s / z^2 + 0.3 (isotropic),
        # extent = ceil(sqrt(q_max) sigma) with q_max = min(2 ln(255 alpha_2d), 8)
        s2 = e["sigma2_px"]
        alpha2d = case["opacity"] * math.sqrt((s2 - 0.3) ** 2 / s2 ** 2)
        qmax = min(2.0 * math.log(255.0 * alpha2d), 8.0)
        assert ex == ey == math.ceil(math.sqrt(qmax) * math.sqrt(s2))
        assert abs(a[5] - alpha2d) < 1e-6
    if e.get("conic_pd"):
        assert a[2] > 0 and a[4] > 0 and a[2] * a[4] - a[3] * a[3] > 0
    if "extent" in e:
        assert ex == ey == e["extent"]


def test_projection_random_splats_match_fp64_twin(oracle_lib):
    """The FP32 restatement against its FP64 twin on a seeded cloud: culling
    decisions agree away from thresholds and attributes agree to f32 rounding."""
    from cindergs_amd import synthetic as syn

    sc = syn.scene_gaussian_ball(20_000)
    cam, comp = syn.view_config1()
    cam_arr = syn.camera_array(cam)
    a32, e32, _ = oracle_lib.project(sc, cam_arr, list(comp.camera_position), 3)
    a64, e64, _ = oracle_lib.project(sc, cam_arr, list(comp.camera_position), 3, fp64=True)
    both = (e32 != 0) & (e64 != 0)
    assert np.mean((e32 != 0) == (e64 != 0)) > 0.999
    rel = np.abs(a32[both, :2] - a64[both, :2]).max()
    assert rel < 1e-2  # pixels
    col = np.abs(a32[both, 6:9] - a64[both, 6:9]).max()
    assert col < 1e-5


# ---------------------------------------------------------------------------
# SH evaluation (sh.rs:143-182) through the projection
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", FIX["sh"], ids=lambda c: c["name"])
def test_sh_known_answers(oracle_lib, case):
    sh = np.zeros((1, 48), np.float32)
    for k, v in case["coeffs"].items():
        sh[0, int(k)] = v
    if case["dir"][2] > 0:  # camera at the origin looking down +Z, splat at z = 3
        view = np.eye(4)
        cam_pos = [0.0, 0.0, 0.0]
    else:  # camera at z = 6 looking down -Z: the view direction is -Z
        view = [[-1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 6], [0, 0, 0, 1]]
        cam_pos = [0.0, 0.0, 6.0]
    cam = _camera24(view, 500.0, 500.0, 250.0, 250.0, 500, 500, 0.2)
    sc = _scene([0.0, 0.0, 3.0], [0.01, 0, 0, 0.01, 0, 0.01], 0.9, sh)
    attrs, ext, surv = oracle_lib.project(sc, cam, cam_pos, sh_degree=case["degree"])
    assert surv == 1
    rgb = attrs[0, 6:9].astype(np.float64)
    assert np.all(np.abs(rgb - np.array(case["expect"])) <= case["tol"] + 1e-12), rgb


# ---------------------------------------------------------------------------
# Stage 3: key layout, binning fixtures, ellipse test, scan, sort
# ---------------------------------------------------------------------------

def test_packed_key_bits(oracle_lib):
    for T, tb, db in FIX["key_bits"]["cases"]:
        assert oracle_lib.packed_key_bits(T) == (tb, db), T
        # tile ids fit the tile field and the key fits 32 bits (or the 64-bit path)
        assert (T - 1) < (1 << tb) or T == 1
        assert db == min(32 - tb, 28)


def _bin(oracle_lib, case, near=0.2):
    """Stage 3 of the restatement on splats given as (centre, radius, depth):
    extents = radius, a near-zero conic so that the ellipse test keeps every
    tile of the extent box (the reference fixtures bin square boxes)."""
    n = len(case["radii"])
    attrs = np.zeros((n, 10), np.float32)
    attrs[:, 0:2] = np.asarray(case["points"], np.float32)
    attrs[:, 2] = 1e-6
    attrs[:, 4] = 1e-6
    attrs[:, 5] = 0.9
    attrs[:, 9] = np.asarray(case["depths"], np.float32)
    r = np.asarray(case["radii"], np.uint32)
    ext = (r | (r << 16)).astype(np.uint32)
    W, H = case["W"], case["H"]
    T = ((W + 15) // 16) * ((H + 15) // 16)
    counts = oracle_lib.count_tiles(attrs, ext, W, H)
    offs, total = oracle_lib.exclusive_scan(counts)
    keys, ids, db = oracle_lib.emit32(attrs, ext, offs, total, W, H, near)
    tb, _ = oracle_lib.packed_key_bits(T)
    sk, sid = oracle_lib.sort_pairs32(keys, ids, min(32, tb + db))
    ranges = oracle_lib.tile_ranges(sk, db, T)
    return dict(total=total, tiles=(sk >> db), ids=sid, ranges=ranges, keys=sk, db=db, T=T,
                depth_q=sk & ((1 << db) - 1))


@pytest.mark.parametrize("case", FIX["binning"], ids=lambda c: c["name"])
def test_binning_fixtures(oracle_lib, case):
    out = _bin(oracle_lib, case)
    e = case["expect"]
    if "total" in e:
        assert out["total"] == e["total"]
        assert out["tiles"].tolist() == e["tiles"]
        assert out["ids"].tolist() == e["ids"]
    for t, (s, en) in e.get("ranges", {}).items():
        assert out["ranges"][int(t)].tolist() == [s, en]
    if "non_empty" in e:
        assert int(np.sum(out["ranges"][:, 1] > out["ranges"][:, 0])) == e["non_empty"]
    # consistency (cpu_reference.rs:496-545): live ids, depth order inside a
    # tile, contiguous total ranges that cover every entry exactly once
    radii = np.asarray(case["radii"])
    assert np.all(radii[out["ids"]] > 0)
    covered = 0
    for t in range(out["T"]):
        s, en = out["ranges"][t]
        if en <= s:
            assert s == en == 0
            continue
        covered += en - s
        assert np.all(out["tiles"][s:en] == t)
        d = np.asarray(case["depths"], np.float32)[out["ids"][s:en]]
        assert np.all(np.diff(d) >= 0)
    assert covered == out["total"]


def test_depth_quantisation_is_monotone(oracle_lib):
    """pack_relative_depth_bits (binning.rs:22-31): a larger depth never gets a
    smaller key inside one tile (checked over 10^5 random depths)."""
    rng = np.random.default_rng(4)
    d = np.sort(rng.uniform(0.21, 500.0, 100_000).astype(np.float32))
    case = {"W": 16, "H": 16, "points": [[8.0, 8.0]] * d.size, "depths": d.tolist(),
            "radii": [1] * d.size}
    out = _bin(oracle_lib, case)
    q_by_id = np.empty(d.size, np.int64)
    q_by_id[out["ids"]] = out["depth_q"]
    assert np.all(np.diff(q_by_id) >= 0)
    assert q_by_id[-1] > q_by_id[0]


def _xorshift_stream(seed):
    s = seed

    def nxt():
        nonlocal s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        return F32((s >> 40) / float(1 << 24))

    return nxt


def test_ellipse_tile_test_has_no_false_negatives(oracle_lib):
    """tiles/mod.rs:302-401 restated: 1000 random ellipses on a 160x128 screen;
    every tile holding a pixel centre that the compositor would blend must be
    emitted, and the tight test must emit < 50% of the isotropic square boxes."""
    W, H = 160, 128
    tx_n, ty_n = W // 16, H // 16
    nxt = _xorshift_stream(0x123456789ABCDEF0)
    rows, iso = [], []
    for _ in range(1000):
        mu_x = F32(-20.0) + nxt() * F32(W + 40.0)
        mu_y = F32(-20.0) + nxt() * F32(H + 40.0)
        ang = nxt() * F32(math.pi)
        l1 = F32(0.5) + nxt() * F32(900.0)
        l2 = F32(0.5) + nxt() * F32(25.0)
        s, c = F32(math.sin(ang)), F32(math.cos(ang))
        sxx = c * c * l1 + s * s * l2 + F32(0.3)
        sxy = s * c * (l1 - l2)
        syy = s * s * l1 + c * c * l2 + F32(0.3)
        alpha = F32(0.01) + nxt() * F32(0.98)
        det = sxx * syy - sxy * sxy
        if det <= F32(1e-6) or alpha <= F32(1.0 / 255.0):
            continue
        sig = F32(math.sqrt(min(max(2.0 * math.log(255.0 * float(alpha)), 0.0), 8.0)))
        inv = F32(1.0) / det
        ex = int(min(max(math.ceil(sig * math.sqrt(sxx)), 1), 1024))
        ey = int(min(max(math.ceil(sig * math.sqrt(syy)), 1), 1024))
        rows.append([mu_x, mu_y, syy * inv, -sxy * inv, sxx * inv, alpha, 0, 0, 0, 1.0, ex, ey])
        iso.append((mu_x, mu_y, math.ceil(3.0 * math.sqrt(max(l1, l2) + 0.3))))
    R = np.asarray(rows, np.float64)
    attrs = np.ascontiguousarray(R[:, :10].astype(np.float32))
    # the O1 box test culls boxes entirely off screen; O2 culls them in projection
    ext = (R[:, 10].astype(np.uint32) | (R[:, 11].astype(np.uint32) << 16)).astype(np.uint32)
    counts = oracle_lib.count_tiles(attrs, ext, W, H)
    offs, total = oracle_lib.exclusive_scan(counts)
    keys, ids, db = oracle_lib.emit32(attrs, ext, offs, total, W, H, 0.2)
    emitted = {(int(i), int(k >> db)) for k, i in zip(keys, ids)}
    px = (np.arange(W, dtype=np.float32) + F32(0.5))[None, :]
    py = (np.arange(H, dtype=np.float32) + F32(0.5))[:, None]
    square_total = 0
    for i in range(attrs.shape[0]):
        mx, my, a, b, c, al = (F32(v) for v in attrs[i, :6])
        dx, dy = px - mx, py - my
        q = F32(-0.5) * (a * dx * dx + F32(2.0) * b * dx * dy + c * dy * dy)
        hit = (q >= F32(-4.0)) & (np.minimum(F32(0.99), al * np.exp(q)) >= F32(1.0 / 255.0))
        hit_t = hit.reshape(ty_n, 16, tx_n, 16).any(axis=(1, 3))
        # in the O2 pipeline projection culls boxes entirely off screen first
        box_on = not (mx + R[i, 10] < 0 or mx - R[i, 10] >= W or my + R[i, 11] < 0
                      or my - R[i, 11] >= H)
        for ty, tx in zip(*np.nonzero(hit_t)):
            if box_on:
                assert (i, ty * tx_n + tx) in emitted, (i, tx, ty)
        cx, cy, r = iso[i]
        if not (cx + r < 0 or cx - r >= W or cy + r < 0 or cy - r >= H):
            x0 = min(max(math.floor((cx - r) / 16), 0), tx_n - 1)
            x1 = min(max(math.floor((cx + r) / 16), 0), tx_n - 1)
            y0 = min(max(math.floor((cy - r) / 16), 0), ty_n - 1)
            y1 = min(max(math.floor((cy + r) / 16), 0), ty_n - 1)
            square_total += (x1 - x0 + 1) * (y1 - y0 + 1)
    assert len(emitted) * 2 < square_total, (len(emitted), square_total)


def test_scan_reference_vectors(oracle_lib):
    for n in FIX["scan"]["sizes"]:
        x = ((np.arange(n, dtype=np.uint64) * 37 + 11) % 7).astype(np.uint32)
        out, tot = oracle_lib.exclusive_scan(x)
        exp = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)[:-1]]).astype(np.uint32)[:n]
        np.testing.assert_array_equal(out, exp)
        assert tot == int(x.sum(dtype=np.uint64)) & 0xFFFFFFFF


def test_radix_reference_vectors(oracle_lib):
    for n in FIX["radix_u32"]["sizes"]:
        i = np.arange(n, dtype=np.uint64)
        keys = (((i * 2654435761) & 0xFFFFFFFF) >> 16 & 0x00FF0FFF).astype(np.uint32)
        vals = i.astype(np.uint32)
        sk, sv = oracle_lib.sort_pairs32(keys, vals)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(sk, keys[order])
        np.testing.assert_array_equal(sv, vals[order])
    # u64 path (gpu_parity_tests.rs:613-661): stable by (tile, depth bits)
    c = FIX["radix_u64"]
    n = c["n"]
    rng = np.random.default_rng(7)
    tiles = rng.integers(0, c["tiles"], n).astype(np.uint32)
    depths = rng.uniform(0.1, 50.0, n).astype(np.float32).view(np.uint32)
    ids = np.arange(n, dtype=np.uint32)
    st, sd, si = oracle_lib.sort64(tiles, depths, ids, c["tile_bits"])
    order = np.lexsort((depths, tiles))  # stable: ties keep index order
    np.testing.assert_array_equal(si, ids[order])
    np.testing.assert_array_equal(st, tiles[order])
    np.testing.assert_array_equal(sd, depths[order])


# ---------------------------------------------------------------------------
# Stage 4 fixtures (rasterizer/cpu_reference.rs:247-358)
# ---------------------------------------------------------------------------

C0 = 0.28209479177387814  # SH band-0 constant, 1 / (2 sqrt(pi))


@pytest.mark.parametrize("case", FIX["compositor"], ids=lambda c: c["name"])
@pytest.mark.parametrize("fp64", [False, True])
def test_compositor_fixtures(oracle_lib, case, fp64):
    W, H = case["W"], case["H"]
    T = ((W + 15) // 16) * ((H + 15) // 16)
    n = len(case["splats"])
    attrs = np.zeros((max(n, 1), 10), np.float32)
    for i, s in enumerate(case["splats"]):
        rgb = [min(max(0.5 + C0 * d, 0.0), 1.0) for d in s["dc"]]
        attrs[i] = [s["mu"][0], s["mu"][1], *s["conic"], s["opacity"], *rgb, s["depth"]]
    ids = np.arange(n, dtype=np.uint32)
    ranges = np.zeros((T, 2), np.uint32)
    if n:
        ranges[0] = [0, n]
    rgba, f = oracle_lib.composite(attrs, ids, ranges, W, H, case["bg"], fp64=fp64, want_float=True)
    e = case["expect"]
    x, y = e["pixel"]
    p = rgba[y, x].astype(np.float64) / 255.0
    if "r_gt" in e:
        assert p[0] > e["r_gt"]
    if "g_lt" in e:
        assert p[1] < e["g_lt"]
    if "rgb" in e:
        assert np.all(np.abs(p[:3] - np.array(e["rgb"])) <= e["tol"] / 255.0 + 1e-9)
        assert rgba.shape == (H, W, 4)


def test_compositor_single_splat_closed_form(oracle_lib):
    """A splat centred on a pixel centre gives alpha = min(0.99, o), colour
    alpha c + (1 - alpha) bg, output alpha 1 - T = alpha, for a sweep of o."""
    for o in [0.01, 0.2, 0.5, 0.75, 0.98, 1.0]:
        attrs = np.array([[4.5, 4.5, 0.5, 0.0, 0.5, o, 0.8, 0.3, 0.1, 1.0]], np.float32)
        ranges = np.zeros((1, 2), np.uint32)
        ranges[0] = [0, 1]
        bg = [0.2, 0.4, 0.6, 1.0]
        rgba = oracle_lib.composite(attrs, np.zeros(1, np.uint32), ranges, 8, 8, bg)
        al = min(0.99, o)
        exp = [round(255 * (al * c + (1 - al) * b)) for c, b in zip([0.8, 0.3, 0.1], bg[:3])]
        exp.append(round(255 * al))
        assert np.all(np.abs(rgba[4, 4].astype(int) - np.array(exp)) <= 1), (o, rgba[4, 4], exp)


# ---------------------------------------------------------------------------
# Stage 1 closed forms (integration_tests.rs:1145-1319, polar.rs:150-210)
# ---------------------------------------------------------------------------

def _box_cage(lo, hi):
    from cindergs_amd.cages import TetMesh

    lo, hi = lo, hi
    V = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                  [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]], np.float32)
    return TetMesh(V, np.array(FIX["box_cage_elements"]["elements"], np.uint32))


def _bind_first(mesh, p):
    """bind_point: first tet in index order with penalty <= 1e-4."""
    p = np.asarray(p, np.float32).reshape(1, 3)
    for t in range(mesh.num_tets):
        w = mesh.barycentric(p, np.array([t]))[0]
        if np.maximum(-w, 0).sum() <= 1e-4:
            return t, w
    raise AssertionError("point outside the cage")


def _rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


@pytest.mark.parametrize("case", FIX["deformation"], ids=lambda c: c["name"])
@pytest.mark.parametrize("fp64", [False, True])
def test_deformation_known_answers(oracle_lib, case, fp64):
    mesh = _box_cage(*case["box"])
    V = mesh.rest_vertices.astype(np.float64)
    if case["map"] == "identity":
        Vd = V
    elif case["map"] == "scale":
        Vd = V * case["s"]
    elif case["map"] == "rot_y":
        Vd = V @ _rot_y(case["angle"]).T
    else:
        Vd = V * np.array([1.0, 1.0, -1.0])
    pos = case.get("pos", [0.1, 0.2, 0.3])
    t, w = _bind_first(mesh, pos)
    cage = TetCage(mesh.elements, mesh.inv_dm, Vd.astype(np.float32), np.array([t], np.uint32),
                   w.reshape(1, 4).astype(np.float32))
    sc = _scene(pos, case.get("cov", [0.04, 0, 0, 0.04, 0, 0.04]), case.get("opacity", 0.5))
    d = oracle_lib.deform(cage, sc, fp64=fp64)
    tol = case.get("tol", 1e-4)
    if case["map"] == "identity":
        np.testing.assert_allclose(d["pos"][0], pos, atol=tol)
        np.testing.assert_allclose(d["cov"][0], case["cov"], atol=tol)
        assert abs(d["opa"][0] - case["opacity"]) < tol
    elif case["map"] == "scale":
        s = case["s"]
        np.testing.assert_allclose(d["pos"][0], np.array(pos) * s, atol=tol)
        for k in (0, 3, 5):
            assert abs(d["cov"][0][k] - case["cov"][k] * s * s) < tol
        s_det = min(max(s ** 3, 0.6), 1.8)
        base = min(max(1.0 - case["opacity"], 1e-4), 0.999)
        exp = min(max(1.0 - base ** (1.0 / s_det), 0.05), 0.99)
        assert abs(d["opa"][0] - exp) < tol
    elif case["map"] == "rot_y":
        rot = _rot_y(case["angle"])
        for ti in range(mesh.num_tets):
            R = np.asarray(d["R"][ti], np.float64).reshape(3, 3).T  # column-major
            assert abs(np.linalg.det(R) - 1.0) < 1e-5
            np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-5)
            np.testing.assert_allclose(R, rot, atol=tol)
    else:
        assert np.all(np.isfinite(d["pos"][0]))
        assert 0.0 <= d["opa"][0] <= 0.99
        assert d["cov"][0][0] >= 1e-4 and d["cov"][0][3] >= 1e-4 and d["cov"][0][5] >= 1e-4


def _polar_of(oracle_lib, J, fp64=False):
    """R (and sigma_max) of one tet whose rest shape is the unit simplex, so
    that D_m = I and the deformation gradient is exactly J."""
    from cindergs_amd.cages import TetMesh

    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    mesh = TetMesh(rest, np.array([[0, 1, 2, 3]], np.uint32))
    J = np.asarray(J, np.float64)
    Vd = np.stack([np.zeros(3), J[:, 0], J[:, 1], J[:, 2]]).astype(np.float32)
    cage = TetCage(mesh.elements, mesh.inv_dm, Vd, np.array([0], np.uint32),
                   np.array([[0.25] * 4], np.float32))
    sc = _scene([0.25, 0.25, 0.25], [0.04, 0, 0, 0.04, 0, 0.04], 0.5)
    d = oracle_lib.deform(cage, sc, fp64=fp64)
    return np.asarray(d["R"][0], np.float64).reshape(3, 3).T, float(d["stretch"][0])


@pytest.mark.parametrize("fp64", [False, True])
def test_polar_decomposition_identities(oracle_lib, fp64):
    R, smax = _polar_of(oracle_lib, np.eye(3), fp64)
    np.testing.assert_allclose(R, np.eye(3), atol=1e-6)
    assert abs(smax - 1.0) < 1e-4
    J = _rot_z(1.2) @ _rot_y(-0.7) @ _rot_x(0.3)
    R, _ = _polar_of(oracle_lib, J, fp64)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-5)
    assert abs(np.linalg.det(R) - 1.0) < 1e-5
    np.testing.assert_allclose(R, J, atol=1e-5)
    rot = _rot_z(math.pi / 4)
    J = rot @ np.diag([2.0, 3.0, 0.5])
    R, smax = _polar_of(oracle_lib, J, fp64)
    np.testing.assert_allclose(R, rot, atol=1e-4)
    assert abs(smax - 3.0) < 1e-3  # largest singular value of J
    R, _ = _polar_of(oracle_lib, np.diag([2.0, 2.0, -2.0]), fp64)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-5)
    assert abs(np.linalg.det(R) - 1.0) < 1e-5
    R, _ = _polar_of(oracle_lib, np.zeros((3, 3)), fp64)
    np.testing.assert_array_equal(R, np.eye(3))


# ---------------------------------------------------------------------------
# Binding (mod.rs:202-323): the oracle's brute-force index-order scan against
# the reference's tests and an independent host restatement (grid search)
# ---------------------------------------------------------------------------

def test_binding_reference_cases(oracle_lib):
    mesh = _box_cage([-1, -1, -1], [1, 1, 1])
    np.testing.assert_array_equal(oracle_lib.tet_inv_dm(mesh.rest_vertices, mesh.elements)
                                  .view(np.uint32), mesh.inv_dm.view(np.uint32))
    t, w = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm,
                                np.array([[0, 0, 0], [0.5, -0.5, 0.5]], np.float32))
    assert np.all(np.abs(w.sum(1) - 1.0) < 1e-5) and np.all(w >= -1e-4)
    mesh = _box_cage([0, 0, 0], [2, 2, 2])
    pt = np.array([[0.7, 1.3, 0.4]], np.float32)
    t, w = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm, pt)
    rec = (w[0][:, None] * mesh.rest_vertices[mesh.elements[t[0]]]).sum(0)
    assert np.all(np.abs(rec - pt[0]) < 1e-4)
    # outside the bounds: unbound; outside the cage without bounds: best penalty
    far = np.array([[5.0, 5.0, 5.0]], np.float32)
    t, w = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm, far,
                                mesh.bounds())
    assert t[0] == 0xFFFFFFFF and np.all(w == 0)
    t, w = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm, far)
    assert t[0] != 0xFFFFFFFF


def test_binding_oracle_equals_grid_search_restatement(oracle_lib):
    """The oracle scans every tet; the host TetMesh.bind_gaussians searches
    only the 27 neighbouring cells of a grid cage.  Same tets and weights."""
    from cindergs_amd.cages import TetMesh

    mesh = TetMesh.jittered_grid([-1, -1, -1], [1, 1, 1], cells=(6, 6, 6))
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.98, 0.98, (8000, 3)).astype(np.float32)
    t, w = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm, p)
    ht, hw = mesh.bind_gaussians(p)
    np.testing.assert_array_equal(t, ht)
    np.testing.assert_array_equal(w.view(np.uint32), hw.view(np.uint32))
    # the first-fit rule: no tet of smaller index accepts the point
    for i in range(0, 8000, 397):
        for tt in range(int(t[i])):
            ww = mesh.barycentric(p[i:i + 1], np.array([tt]))[0]
            assert np.maximum(-ww, 0).sum() > 1e-4


# ---------------------------------------------------------------------------
# Mean value coordinates (mvc.rs:366-541) against the reference's tests and
# the identities that define them
# ---------------------------------------------------------------------------

def _mvc_cage(kind, a, b):
    m = FIX["mvc_cages"]
    if kind == "box":
        lo, hi = a, b
        V = [[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
             [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
             [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]]
        return np.array(V, np.float32), np.array(m["box_triangles"], np.uint32)
    c, r = np.array(a, np.float32), np.float32(max(abs(b), 1e-4))
    if kind == "octahedron":
        d = [[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r], [0, 0, -r]]
        return (c + np.array(d, np.float32)).astype(np.float32), \
            np.array(m["octahedron_triangles"], np.uint32)
    phi = (1.0 + math.sqrt(5.0)) * 0.5
    raw = np.array([[-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0], [0, -1, phi],
                    [0, 1, phi], [0, -1, -phi], [0, 1, -phi], [phi, 0, -1], [phi, 0, 1],
                    [-phi, 0, -1], [-phi, 0, 1]], np.float32)
    raw = raw / np.linalg.norm(raw, axis=1, keepdims=True)
    return (c + raw * r).astype(np.float32), np.array(m["icosahedron_triangles"], np.uint32)


def test_mvc_presets_and_icosphere_are_closed():
    from cindergs_amd.mvc import PolyCage, icosahedron, icosphere

    V, T = _mvc_cage("box", [-1.0] * 3, [1.0] * 3)
    box = PolyCage(V, T)
    assert box.is_closed_manifold() and abs(box.signed_volume() - 8.0) < 1e-4
    for kind, a, b in (("octahedron", [0, 0, 0], 1.5), ("icosahedron", [0, 0, 0], 2.0)):
        c = PolyCage(*_mvc_cage(kind, a, b))
        assert c.is_closed_manifold() and c.signed_volume() > 0
    ico = icosahedron()
    assert ico.is_closed_manifold() and ico.signed_volume() > 0
    sph = icosphere()
    assert sph.num_vertices == 642 and sph.triangles.shape[0] == 1280
    assert sph.is_closed_manifold() and sph.signed_volume() > 0


@pytest.mark.parametrize("fp64", [False, True])
def test_mvc_reference_reproduction(oracle_lib, fp64):
    case = FIX["mvc"][1]
    for kind, a, b in case["cages"]:
        V, T = _mvc_cage(kind, a, b)
        w, _, inside = oracle_lib.mvc_bind(V, T, np.array([case["point"]], np.float32), fp64)
        assert inside[0]
        assert abs(w[0].sum() - 1.0) < case["sum_tol"]
        assert np.linalg.norm(w[0] @ V.astype(np.float64) - case["point"]) < case["reproduction_tol"]
        if kind != "box":  # convex cages: positive coordinates
            assert w[0].min() > -1e-6
    c = FIX["mvc"][2]
    V, T = _mvc_cage("octahedron", *c["octahedron"])
    w = oracle_lib.mvc_weights(V, T, np.array([c["vertex"], c["face_point"]], np.float32), fp64)
    assert abs(w[0][c["vertex_index"]] - 1.0) < 1e-6
    assert np.linalg.norm(w[1] @ V.astype(np.float64) - c["face_point"]) < c["tol"]


@pytest.mark.parametrize("fp64", [False, True])
def test_mvc_gradients_match_local_jacobian(oracle_lib, fp64):
    """mvc.rs:885-907: sum_j v'_j grad(w_j)^T equals the central difference of
    the deformed position; for the undeformed cage both are the identity
    (linear precision)."""
    c = FIX["mvc"][3]
    V, T = _mvc_cage("octahedron", *c["octahedron"])
    pts = np.array(c["points"], np.float32)
    w, g, inside = oracle_lib.mvc_bind(V, T, pts, fp64)
    assert inside.all()
    Vd = V.astype(np.float64).copy()
    for vi, ax, dv in c["deform"]:
        Vd[vi, ax] += dv
    lo, hi = V.min(0), V.max(0)
    h = min(max(np.float32(np.linalg.norm(hi - lo)) * np.float32(1e-3), 1e-5), 1e-2)
    for i, p in enumerate(pts):
        F = np.einsum("jk,jl->lk", g[i], Vd)  # column k = sum_j v'_j g_jk
        cols = []
        for ax in range(3):
            pp, pm = p.copy(), p.copy()
            pp[ax] += np.float32(h)
            pm[ax] -= np.float32(h)
            wp, _, _ = oracle_lib.mvc_bind(V, T, pp[None], fp64, gradients=False)
            wm, _, _ = oracle_lib.mvc_bind(V, T, pm[None], fp64, gradients=False)
            cols.append((wp[0] @ Vd - wm[0] @ Vd) * (0.5 / h))
        assert np.abs(F - np.stack(cols, 1)).max() < c["tol"]
        Fi = np.einsum("jk,jl->lk", g[i], V.astype(np.float64))
        assert np.abs(Fi - np.eye(3)).max() < (1e-3 if not fp64 else 1e-6)


def test_mvc_fp32_matches_fp64_on_the_icosphere(oracle_lib):
    from cindergs_amd.mvc import icosphere

    cage = icosphere()
    rng = np.random.default_rng(9)
    d = rng.normal(size=(64, 3))
    p = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0, 0.95, (64, 1))
         ).astype(np.float32)  # inside the radius 1.2 - 0.15 ball
    w32, _, in32 = oracle_lib.mvc_bind(cage.vertices, cage.triangles, p, False, gradients=False)
    w64, _, in64 = oracle_lib.mvc_bind(cage.vertices, cage.triangles, p, True, gradients=False)
    assert in32.all() and in64.all()
    # near the surface the small angles lose relative precision in f32: the
    # weights agree to 1e-4 absolute, the reproduced positions to 1e-5 of the
    # cage diagonal
    assert np.abs(w32 - w64).max() < 1e-4
    diag = cage.diagonal()
    assert np.abs(w32.astype(np.float64) @ cage.vertices.astype(np.float64) - p).max() < 1e-5 * diag
    assert np.abs(w64.sum(1) - 1).max() < 1e-12
    assert np.abs(w64 @ cage.vertices.astype(np.float64) - p).max() < 1e-12
    far = np.array([[3.0, 0, 0], [0, -2.5, 0.1]], np.float32)
    assert np.all(np.abs(oracle_lib.mvc_winding(cage.vertices, cage.triangles, far, True)) < 1e-6)


# ---------------------------------------------------------------------------
# The surrogate network's CPU restatement: its backward pass against central
# differences of its own loss (float64, no bf16 rounding)
# ---------------------------------------------------------------------------

def test_mlp_oracle_backward_matches_finite_differences():
    import mlp_oracle as mo

    rng = np.random.default_rng(11)
    n_out = 20
    rows = [256, 256, 256, 256, 32]
    cols = [64, 256, 256, 256, 256]
    offs = list(np.cumsum([0] + [r * (c + 32) for r, c in zip(rows, cols)])[:-1])
    params = np.zeros(sum(r * (c + 32) for r, c in zip(rows, cols)))
    for r, c, o in zip(rows, cols, offs):
        blk = params[o:o + r * (c + 32)].reshape(r, c + 32)
        blk[:, :c + 1] = rng.normal(0, 1.0 / math.sqrt(c), (r, c + 1))
        if r == 32:
            blk[n_out:] = 0.0
    fourier = rng.normal(0, 1.0, (32, 3))
    x = rng.uniform(-1, 1, (16, 3))
    t = rng.dirichlet(np.ones(n_out), 16)
    loss, grad = mo.loss_and_grad(params, fourier, rows, cols, offs, x, t, n_out, use_bf16=False)
    for l, (r, c, o) in enumerate(zip(rows, cols, offs)):
        for (i, j) in [(0, 0), (r // 2 if r > n_out else n_out - 1, c // 3), (1, c)]:  # c = bias
            k = o + i * (c + 32) + j
            eps = 1e-6
            p1, p2 = params.copy(), params.copy()
            p1[k] += eps
            p2[k] -= eps
            l1, _ = mo.loss_and_grad(p1, fourier, rows, cols, offs, x, t, n_out, use_bf16=False)
            l2, _ = mo.loss_and_grad(p2, fourier, rows, cols, offs, x, t, n_out, use_bf16=False)
            fd = (l1 - l2) / (2 * eps)
            assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, i, j, fd, grad[k])
    # the normalised weights sum to one whatever the network
    w = mo.weights(params, fourier, rows, cols, offs, x, n_out, use_bf16=False)
    assert np.abs(w.sum(1) - 1).max() < 1e-12
