"""Exact mean value coordinates on the GPU (cgs_mvc_*) against the oracle's
FP64 restatement of PolyCageMesh (src/deformation/mvc.rs:326-541).

Bars: the inside mask identical to the FP64 oracle's; weights, reproduced
positions (sum_j w_j v_j = x, linear precision), gradients and the rest
Jacobian sum_j v_j grad(w_j)^T = I each within twice the error that the
reference algorithm itself makes in f32 (the oracle's FP32 restatement,
measured on the same points against FP64: about 1e-4 on weights near the
surface, 7e-5 on positions, 1.6e-2 on the central-difference Jacobian);
deformed positions and Jacobians equal to float64 sums over the same weights
to 1e-5 relative."""
import json
import os

import numpy as np
import pytest

from cage_presets import preset as _preset
from cindergs_amd.mvc import MvcGpu, icosphere

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


def _ball_points(n, r, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True) * r * rng.uniform(0, 1, (n, 1)) ** (1 / 3)
            ).astype(np.float32)


def test_icosphere_bind_matches_fp64_oracle(gpu, oracle_lib):
    cage = icosphere()
    inside = _ball_points(1500, 1.0, 1)
    outside = (_ball_points(200, 1.0, 2) * 0.1 + np.array([0, 0, 2.5], np.float32)).astype(np.float32)
    p = np.concatenate([inside, outside]).astype(np.float32)
    print(f"[case] mvc icosphere bind n={p.shape[0]}", flush=True)
    mg = MvcGpu(gpu, cage)
    dp = gpu.to_device(p)
    w, g, m = mg.bind(dp, p.shape[0])
    W, G, M = w.numpy(), g.numpy(), m.numpy().astype(bool)
    rw, rg, rm = oracle_lib.mvc_bind(cage.vertices, cage.triangles, p, fp64=True)
    fw, fg, _ = oracle_lib.mvc_bind(cage.vertices, cage.triangles, p, fp64=False)
    np.testing.assert_array_equal(M, rm)
    assert M[:1500].all() and not M[1500:].any()
    V = cage.vertices.astype(np.float64)

    def errors(w, g):
        return (np.abs(w[M] - rw[M]).max(),
                np.abs(w[M].astype(np.float64) @ V - p[M]).max(),
                np.abs(np.einsum("njk,jl->nlk", g[M].astype(np.float64), V) - np.eye(3)).max(),
                np.abs(g[M] - rg[M]).max())

    gpu_e, f32_e = errors(W, G), errors(fw, fg)
    print("[case] mvc errors vs FP64 (weights, positions, |F-I|, gradients): gpu "
          + " ".join(f"{e:.2e}" for e in gpu_e) + " | reference in f32 "
          + " ".join(f"{e:.2e}" for e in f32_e), flush=True)
    for e_gpu, e_f32 in zip(gpu_e, f32_e):
        assert e_gpu <= 2.0 * e_f32 + 1e-7
    assert gpu_e[1] < 5e-5 * cage.diagonal()
    assert np.all(W[~M] == 0) and np.all(G[~M] == 0)
    for b in (dp, w, g, m):
        b.free()


def test_reference_mvc_cases_on_gpu(gpu):
    case = FIX["mvc"][1]
    for kind, a, b in case["cages"]:
        cage = _preset(kind, a, b)
        mg = MvcGpu(gpu, cage)
        dp = gpu.to_device(np.array([case["point"]], np.float32))
        w, g, m = mg.bind(dp, 1)
        W = w.numpy()[0]
        assert m.numpy()[0] == 1
        assert abs(W.sum() - 1.0) < case["sum_tol"]
        assert np.linalg.norm(W @ cage.vertices.astype(np.float64) - case["point"]) < \
            case["reproduction_tol"]
    # precomputed gradients against the central difference of deformed positions
    c = FIX["mvc"][3]
    cage = _preset("octahedron", *c["octahedron"])
    pts = np.array(c["points"], np.float32)
    mg = MvcGpu(gpu, cage)
    dp = gpu.to_device(pts)
    w, g, m = mg.bind(dp, 3)
    Vd = cage.vertices.copy()
    for vi, ax, dv in c["deform"]:
        Vd[vi, ax] += np.float32(dv)
    pos, F = mg.deform(w, g, m, dp, Vd, 3)
    Fm = F.numpy().reshape(3, 3, 3).transpose(0, 2, 1)  # column-major -> F[i][row][col]
    lo, hi = cage.vertices.min(0), cage.vertices.max(0)
    h = np.float32(min(max(np.float32(np.linalg.norm(hi - lo)) * np.float32(1e-3), 1e-5), 1e-2))
    for i in range(3):
        cols = []
        for ax in range(3):
            pp, pm = pts[i].copy(), pts[i].copy()
            pp[ax] += h
            pm[ax] -= h
            dpp = gpu.to_device(np.stack([pp, pm]))
            ww, _, mm = mg.bind(dpp, 2, gradients=False)
            Wd = ww.numpy().astype(np.float64) @ Vd.astype(np.float64)
            cols.append((Wd[0] - Wd[1]) * (0.5 / float(h)))
        assert np.abs(Fm[i] - np.stack(cols, 1)).max() < 1e-4


def test_deform_kernel_matches_float64_sums(gpu):
    cage = icosphere()
    p = _ball_points(512, 1.0, 3)
    p = np.concatenate([p, np.array([[0, 0, 3.0]], np.float32)])
    mg = MvcGpu(gpu, cage)
    dp = gpu.to_device(p)
    w, g, m = mg.bind(dp, p.shape[0])
    rng = np.random.default_rng(4)
    Vd = (cage.vertices + rng.uniform(-0.1, 0.1, cage.vertices.shape)).astype(np.float32)
    pos, F = mg.deform(w, g, m, dp, Vd, p.shape[0])
    W, G, M = w.numpy().astype(np.float64), g.numpy().astype(np.float64), m.numpy().astype(bool)
    P = pos.numpy()
    ref = W @ Vd.astype(np.float64)
    assert np.allclose(P[M], ref[M], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(P[~M], p[~M])
    # F = sum_j v'_j g_j^T, column-major [n, 9]
    Fd = F.numpy().reshape(-1, 3, 3).transpose(0, 2, 1)
    Fref = np.einsum("jr,njc->nrc", Vd.astype(np.float64), G.reshape(-1, cage.num_vertices, 3))
    # f32 sums of 642 terms: within 1e-6 of the sum of magnitudes
    mag = np.einsum("jr,njc->nrc", np.abs(Vd.astype(np.float64)),
                    np.abs(G.reshape(-1, cage.num_vertices, 3)))
    assert np.all(np.abs(Fd[M] - Fref[M]) <= 1e-6 * mag[M] + 1e-7)
    np.testing.assert_array_equal(Fd[~M], np.broadcast_to(np.eye(3, dtype=np.float32),
                                                          Fd[~M].shape))
    # positions alone (no Jacobian output) give the same positions
    pos2, _ = mg.deform(w, g, m, dp, Vd, p.shape[0], jacobians=False)
    np.testing.assert_array_equal(pos2.numpy(), P)
    Fr = np.einsum("njk,jl->nkl", G, Vd.astype(np.float64)).reshape(-1, 9)  # column-major
    assert np.allclose(F.numpy()[M], Fr[M], rtol=1e-4, atol=1e-5)
    np.testing.assert_array_equal(F.numpy()[~M][0], np.eye(3).reshape(-1).astype(np.float32))
