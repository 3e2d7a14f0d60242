"""Pins of the oracle's Stage 1 split / 3D-mip branch and of its 64-bit key
emission (CPU only).

The split branch restates deform_and_split_compact_kernel
(src/render/kernels/deformation.rs:1066-1211 of the reference): a splat whose
tet stretches more than tau is replaced by two children at p -/+ 1/2 sqrt(l0) u
(l0, u: the major eigenpair of the transported covariance M = J S J^T) that
share the covariance M - 3/4 l0 u u^T, and the 3D mip filter (1020-1060,
1180-1196) adds var to the diagonal and scales the opacity by
sqrt(det C / det(C + var I)).  Here the cage is deformed by a known linear map
A, so J = A exactly and every expected value follows from float64 linear
algebra (numpy's eigh, det), independent of the oracle's Cardano solve: a
dropped square root, a wrong 3/4, a transposed J or an eigenvector of the
wrong eigenvalue fails these checks.

The 64-bit emission (emit_tile_keys_u64_kernel, binning.rs:254-328) is run on
the reference's binning fixtures (tiles/cpu_reference.rs:409-556) through the
64-bit sort and range path, which must give the fixtures' tiles, ids and
ranges exactly.
"""
import math

import numpy as np
import pytest

from cindergs_amd.pipeline import GaussianSceneSoa, TetCage
from test_oracle_golden import FIX, _bind_first, _box_cage


def _rand_rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _sym6(m):
    return np.array([m[0, 0], m[0, 1], m[0, 2], m[1, 1], m[1, 2], m[2, 2]])


def _mat(c6):
    c6 = np.asarray(c6, np.float64)
    return np.array([[c6[0], c6[1], c6[2]], [c6[1], c6[3], c6[4]], [c6[2], c6[4], c6[5]]])


def _f(x):
    """a literal of the reference kernel (an f32 constant) as a float64"""
    return float(np.float32(x))


def _density_alpha(alpha, det_a):
    s_det = min(max(det_a, _f(0.6)), _f(1.8))
    base = min(max(1.0 - alpha, _f(1e-4)), _f(0.999))
    return min(max(1.0 - base ** (1.0 / s_det), _f(0.05)), _f(0.99))


def _mip(C, alpha, var):
    det_o = max(np.linalg.det(C), _f(1e-24))
    Cr = C + _f(var) * np.eye(3)
    det_r = max(np.linalg.det(Cr), det_o)
    comp = math.sqrt(min(max(det_o / det_r, 0.0), 1.0))
    return Cr, min(max(alpha * comp, 0.0), 1.0)


def _dyadic_map(rng, diag):
    """A linear map with entries on a 1/16 grid: exact in f32, so the deformed
    box-cage vertices (+-2) and J = A carry no rounding."""
    while True:
        A = np.diag(np.asarray(diag, np.float64)) + rng.integers(-4, 5, (3, 3)) / 16.0
        if np.linalg.det(A) > 0.2:
            return A


def _setup(rng, n, A, unbound=0):
    """n splats inside a [-2, 2]^3 box cage deformed by x -> A x; the last
    `unbound` splats carry no tet."""
    mesh = _box_cage([-2, -2, -2], [2, 2, 2])
    pos = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    covs, tets, ws = [], [], []
    for i in range(n):
        Rr = _rand_rot(rng)
        s = rng.uniform(0.05, 0.3, 3)
        covs.append(_sym6(Rr @ np.diag(s * s) @ Rr.T))
        t, w = _bind_first(mesh, pos[i])
        tets.append(t if i < n - unbound else 0xFFFFFFFF)
        ws.append(w if i < n - unbound else np.zeros(4))
    V = mesh.rest_vertices.astype(np.float64)
    Vd = (V @ np.asarray(A, np.float64).T).astype(np.float32)
    cage = TetCage(mesh.elements, mesh.inv_dm, Vd, np.asarray(tets, np.uint32),
                   np.asarray(ws, np.float32))
    opa = rng.uniform(0.1, 0.95, n).astype(np.float32)
    scene = GaussianSceneSoa(pos, np.asarray(covs, np.float32), opa,
                             np.zeros((n, 48), np.float32))
    return mesh, cage, scene


REST_BOX = _box_cage([-2, -2, -2], [2, 2, 2]).rest_vertices.astype(np.float64)


def _expected(scene, cage, A, tau, mip, stretch):
    """float64 restatement from A: records in output order (parents, then the
    second children in parent order)."""
    A = np.asarray(A, np.float64)
    det_a = float(np.linalg.det(A))
    n = scene.count
    first, second = [], []
    for i in range(n):
        x = scene.positions[i].astype(np.float64)
        S = _mat(scene.covariances_3d[i])
        alpha = float(scene.opacities[i])
        if cage.bind_tet[i] == 0xFFFFFFFF:
            C, a = (S, alpha) if mip <= 0 else _mip(S, alpha, mip)
            first.append((x, C, a))
            continue
        # the position the f32 barycentric weights reconstruct on the rest cage
        w = cage.bind_weights[i].astype(np.float64)
        e = cage.elements[cage.bind_tet[i]]
        xr = sum(w[k] * REST_BOX[e[k]] for k in range(4))
        p = A @ xr
        M = A @ S @ A.T
        for k in range(3):
            M[k, k] = max(M[k, k], _f(1e-4))
        a = _density_alpha(alpha, det_a)
        if not (tau > 0 and stretch > tau):
            C, a2 = (M, a) if mip <= 0 else _mip(M, a, mip)
            first.append((p, C, a2))
            continue
        lam, vec = np.linalg.eigh(M)
        l0, u = lam[-1], vec[:, -1]
        d = 0.5 * math.sqrt(l0) * u
        C = M - 0.75 * l0 * np.outer(u, u)
        for k in range(3):
            C[k, k] = max(C[k, k], _f(1e-8))
        if mip > 0:
            C, a = _mip(C, a, mip)
        first.append((p - d, C, a, p, u, l0, M))
        second.append((p + d, C, a, i))
    return first, second


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("mip", [0.0, 2e-3])
def test_split_branch_matches_closed_form(oracle_lib, fp64, mip):
    rng = np.random.default_rng(11 if fp64 else 12)
    A = _dyadic_map(rng, [2.25, 1.125, 0.875])
    n = 24
    mesh, cage, scene = _setup(rng, n, A, unbound=3)
    tau = 1.5
    d = oracle_lib.deform(cage, scene, tau=tau, mip=mip, fp64=fp64)
    rtol = 1e-10 if fp64 else 2e-5
    # every tet of the linear map: J = A and sigma_max = the largest singular value
    sig = float(np.linalg.svd(A, compute_uv=False)[0])
    for t in range(mesh.num_tets):
        J = np.asarray(d["J"][t], np.float64).reshape(3, 3).T
        np.testing.assert_allclose(J, A, rtol=0, atol=4 * rtol * 2.2)
        assert abs(float(d["stretch"][t]) - sig) <= 1e-4 * sig
    assert sig > tau
    first, second = _expected(scene, cage, A, tau, mip, sig)
    assert d["splits"] == len(second) == n - 3
    # second children in parent order, with (parent, tet) metadata
    np.testing.assert_array_equal(d["split_meta"][:, 0], [s[3] for s in second])
    np.testing.assert_array_equal(d["split_meta"][:, 1], cage.bind_tet[[s[3] for s in second]])
    for i, rec in enumerate(first):
        got_p = np.asarray(d["pos"][i], np.float64)
        got_c = _mat(d["cov"][i])
        scale_c = np.abs(rec[1]).max()
        exp_p = rec[0]
        if len(rec) > 3:
            # u is defined up to sign: the first child sits at p - d or p + d
            alt = 2.0 * rec[3] - rec[0]
            if np.abs(got_p - alt).max() < np.abs(got_p - exp_p).max():
                exp_p = alt
        assert np.abs(got_p - exp_p).max() <= rtol * max(1.0, np.abs(exp_p).max()) * 4, i
        assert np.abs(got_c - rec[1]).max() <= rtol * scale_c * 8, i
        assert abs(float(d["opa"][i]) - rec[2]) <= rtol * 8, i
        if len(rec) > 3:
            # invariants of the split: M u = l0 u, children straddle the parent
            # at +-1/2 sqrt(l0) u, the child variance along u is l0 / 4
            p, u, l0, M = rec[3], rec[4], rec[5], rec[6]
            np.testing.assert_allclose(M @ u, l0 * u, atol=1e-12 * l0)
            j = [s[3] for s in second].index(i)
            twin = np.asarray(d["pos"][n + j], np.float64)
            assert np.abs(0.5 * (got_p + twin) - p).max() <= rtol * 8 * max(1.0, np.abs(p).max())
            off = 0.5 * (twin - got_p)
            assert abs(np.linalg.norm(off) - 0.5 * math.sqrt(l0)) <= rtol * 8
            assert abs(abs(off @ u) - 0.5 * math.sqrt(l0)) <= rtol * 8
            if mip == 0.0:
                assert abs(u @ got_c @ u - 0.25 * l0) <= rtol * 8 * l0
    for j, rec in enumerate(second):
        got_c = _mat(d["cov"][n + j])
        np.testing.assert_array_equal(d["cov"][n + j], d["cov"][rec[3]])
        assert abs(float(d["opa"][n + j]) - rec[2]) <= rtol * 8
        assert np.abs(got_c - rec[1]).max() <= rtol * np.abs(rec[1]).max() * 8


@pytest.mark.parametrize("fp64", [False, True])
def test_split_axis_of_a_diagonal_covariance(oracle_lib, fp64):
    """p1 <= 1e-10 (deformation.rs:1069-1079): a diagonal M splits along the
    axis of its largest diagonal entry, first match in x, y, z order."""
    rng = np.random.default_rng(3)
    for axis in range(3):
        scale = [1.0, 1.0, 1.0]
        scale[axis] = 2.0
        A = np.diag(scale)
        mesh, cage, scene = _setup(rng, 2, A)
        scene.covariances_3d[:] = np.array([0.01, 0, 0, 0.01, 0, 0.01], np.float32)
        d = oracle_lib.deform(cage, scene, tau=1.5, fp64=fp64)
        assert d["splits"] == 2
        for i in range(2):
            off = np.asarray(d["pos"][2 + i], np.float64) - np.asarray(d["pos"][i], np.float64)
            expect = np.zeros(3)
            expect[axis] = math.sqrt(0.04)  # 2 * 1/2 sqrt(l0), l0 = 4 * 0.01
            np.testing.assert_allclose(off, expect, atol=1e-6)
            c = _mat(d["cov"][i])
            assert abs(c[axis, axis] - 0.01) < 1e-7  # 0.04 - 3/4 * 0.04


@pytest.mark.parametrize("fp64", [False, True])
def test_mip_filter_without_split_matches_closed_form(oracle_lib, fp64):
    rng = np.random.default_rng(5)
    A = _dyadic_map(rng, [1.25, 1.0, 0.75])
    mesh, cage, scene = _setup(rng, 16, A, unbound=4)
    mip = 5e-3
    d = oracle_lib.deform(cage, scene, tau=0.0, mip=mip, fp64=fp64)
    assert d["splits"] == 0
    first, _ = _expected(scene, cage, A, 0.0, mip, 0.0)
    rtol = 1e-10 if fp64 else 2e-5
    for i, rec in enumerate(first):
        got_c = _mat(d["cov"][i])
        assert np.abs(got_c - rec[1]).max() <= rtol * 8 * np.abs(rec[1]).max()
        assert abs(float(d["opa"][i]) - rec[2]) <= rtol * 8
        # attenuation: det(C)/det(C + var I) < 1 lowers the opacity
        assert float(d["opa"][i]) <= (float(scene.opacities[i]) if i >= 12 else 0.99) + 1e-6


def _bin64(oracle_lib, case):
    near = 0.2
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
    tiles, depths, ids = oracle_lib.emit64(attrs, ext, offs, total, W, H)
    tb, _ = oracle_lib.packed_key_bits(T)
    st, sd, sid = oracle_lib.sort64(tiles, depths, ids, tb)
    ranges = oracle_lib.tile_ranges(st, 0, T)
    return dict(total=total, tiles=st, depths=sd, ids=sid, ranges=ranges, T=T, attrs=attrs,
                raw_tiles=tiles, raw_depths=depths, raw_ids=ids, near=near)


@pytest.mark.parametrize("case", FIX["binning"], ids=lambda c: c["name"])
def test_emit64_on_binning_fixtures(oracle_lib, case):
    out = _bin64(oracle_lib, case)
    e = case["expect"]
    if "total" in e:
        assert out["total"] == e["total"]
        assert out["tiles"].tolist() == e["tiles"]
        assert out["ids"].tolist() == e["ids"]
    for t, (s, en) in e.get("ranges", {}).items():
        assert out["ranges"][int(t)].tolist() == [s, en]
    if "non_empty" in e:
        assert int(np.sum(out["ranges"][:, 1] > out["ranges"][:, 0])) == e["non_empty"]
    # emission: splat-major, tiles row-major, the depth word is the raw f32
    # bit pattern of max(tz, 0) (binning.rs:300-318)
    ids = out["raw_ids"]
    assert np.all(np.diff(ids.astype(np.int64)) >= 0)
    for k in range(1, ids.size):
        if ids[k] == ids[k - 1]:
            assert out["raw_tiles"][k] > out["raw_tiles"][k - 1]
    tz = np.maximum(out["attrs"][ids, 9], np.float32(0.0)).astype(np.float32)
    np.testing.assert_array_equal(out["raw_depths"], tz.view(np.uint32))
    # the 32-bit path gives the same tiles and ids wherever its quantised
    # depths do not tie
    from test_oracle_golden import _bin

    o32 = _bin(oracle_lib, case)
    np.testing.assert_array_equal(o32["tiles"], out["tiles"])
    if np.unique(np.asarray(case["depths"], np.float32)).size == len(case["depths"]):
        np.testing.assert_array_equal(o32["ids"], out["ids"])


@pytest.mark.parametrize("mip", [1e-4, 2e-3])
def test_stage2_mip_filter_equals_projecting_the_regularised_splat(oracle_lib, mip):
    """project.rs:1114-1133 without cages: projecting (Sigma, alpha) with the
    3D mip filter must equal projecting (Sigma + var I, alpha * sqrt(det Sigma /
    det(Sigma + var I))) without it, the factor computed here in float64; a
    splat the factor pushes below min_opacity is culled.  Checked on the FP64
    twin, whose unfiltered projection the closed forms above pin."""
    from cindergs_amd import synthetic as syn

    sc = syn.scene_gaussian_ball(5_000, k=31)
    cam, comp = syn.view_config1()
    cam_arr = syn.camera_array(cam)
    a_m, e_m, _ = oracle_lib.project(sc, cam_arr, list(comp.camera_position), 3, mip=mip,
                                     fp64=True)
    C = sc.covariances_3d.astype(np.float64)
    full = np.stack([_mat(c) for c in C])
    det_o = np.maximum(np.linalg.det(full), _f(1e-24))
    reg = full + _f(mip) * np.eye(3)
    det_r = np.maximum(np.linalg.det(reg), det_o)
    comp_f = np.sqrt(np.clip(det_o / det_r, 0.0, 1.0))
    alpha = np.clip(sc.opacities.astype(np.float64) * comp_f, 0.0, 1.0)
    assert comp_f.max() < 1.0 and comp_f.min() < 0.9  # the filter bites
    C2 = C.copy()
    C2[:, [0, 3, 5]] += _f(mip)
    sc2 = GaussianSceneSoa(sc.positions, C2.astype(np.float32), alpha.astype(np.float32),
                           sc.sh_coeffs)
    a_r, e_r, _ = oracle_lib.project(sc2, cam_arr, list(comp.camera_position), 3, fp64=True)
    culled_by_filter = (sc.opacities >= 1.0 / 255.0) & (alpha < 1.0 / 255.0)
    assert not np.any(e_m[culled_by_filter])
    # the f32 rounding of the regularised inputs moves a few extents by one px
    both = (e_m != 0) & (e_r != 0)
    assert np.mean((e_m != 0) == (e_r != 0)) > 0.999
    assert np.mean(e_m[both] == e_r[both]) > 0.99
    np.testing.assert_array_equal(a_m[both, :2], a_r[both, :2])   # means
    np.testing.assert_array_equal(a_m[both, 6:10], a_r[both, 6:10])  # colour, depth
    rel = np.abs(a_m[both, 2:6] - a_r[both, 2:6]) / np.maximum(np.abs(a_r[both, 2:6]), 1e-12)
    assert np.max(rel[:, [0, 2, 3]]) < 1e-5  # conic a, c and alpha_2d
    b_scale = np.maximum(np.abs(a_r[both, 2]), np.abs(a_r[both, 4]))
    assert np.max(np.abs(a_m[both, 3] - a_r[both, 3]) / b_scale) < 1e-5
    # the FP32 restatement agrees with its FP64 twin under the filter
    a32, e32, _ = oracle_lib.project(sc, cam_arr, list(comp.camera_position), 3, mip=mip)
    ok = (e32 != 0) & (e_m != 0)
    assert np.mean((e32 != 0) == (e_m != 0)) > 0.999
    assert np.abs(a32[ok, 5] - a_m[ok, 5]).max() < 1e-5
