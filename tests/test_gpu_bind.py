"""Splat-to-tet binding on the GPU (cgs_bind_tets) against the oracle's
restatement of TetMesh::bind_point (src/deformation/mod.rs:267-310): tet
indices and weights bit-exact, with and without bounds, including points on
shared faces and vertices (where index order decides) and points outside the
cage (the smallest-penalty fallback); plus the reference's binding tests
(mod.rs:523-600)."""
import json
import os

import numpy as np
import pytest

from cindergs_amd.cages import TetMesh, auto_bind_gpu

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))
UNBOUND = 0xFFFFFFFF


def _box(lo, hi):
    V = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
                  [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]], np.float32)
    return TetMesh(V, np.array(FIX["box_cage_elements"]["elements"], np.uint32))


def _points(mesh, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = mesh.rest_vertices.min(0), mesh.rest_vertices.max(0)
    ext = hi - lo
    inside = rng.uniform(lo, hi, (n, 3))
    outside = rng.uniform(lo - 0.3 * ext, hi + 0.3 * ext, (n // 10, 3))
    verts = mesh.rest_vertices[rng.integers(0, mesh.num_vertices, n // 20)]
    e = mesh.elements[rng.integers(0, mesh.num_tets, n // 20)]
    X = mesh.rest_vertices
    faces = (X[e[:, 0]] + X[e[:, 1]] + X[e[:, 2]]) / np.float32(3.0)
    edges = (X[e[:, 0]] + X[e[:, 3]]) * np.float32(0.5)
    return np.ascontiguousarray(np.concatenate([inside, outside, verts, faces, edges]),
                                dtype=np.float32)


@pytest.mark.parametrize("with_bounds", [False, True])
def test_grid_cage_binding_matches_oracle(gpu, oracle_lib, with_bounds):
    mesh = TetMesh.jittered_grid([-2.0, -1.5, -1.0], [2.0, 1.5, 1.0])  # 6000 tets
    p = _points(mesh, 60_000, 5)
    bounds = mesh.bounds() if with_bounds else None
    print(f"[case] bind grid cage: {mesh.num_tets} tets, {p.shape[0]} points, bounds={with_bounds}",
          flush=True)
    t, w = mesh.bind_gaussians_gpu(gpu, p, bounds)
    rt, rw = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm, p, bounds)
    np.testing.assert_array_equal(t, rt)
    np.testing.assert_array_equal(w.view(np.uint32), rw.view(np.uint32))
    assert (t == UNBOUND).any() == with_bounds


def test_degenerate_and_single_tet_meshes(gpu, oracle_lib):
    # a flat tet (identity inverse) before a regular one, and a lone tet
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.5, 0], [0, 0, 1]], np.float32)
    E = np.array([[0, 1, 2, 3], [0, 1, 2, 4]], np.uint32)
    mesh = TetMesh(V, E)
    rng = np.random.default_rng(2)
    p = rng.uniform(-0.5, 1.5, (5000, 3)).astype(np.float32)
    for elems in (E, E[1:]):
        m = TetMesh(V, elems)
        t, w = m.bind_gaussians_gpu(gpu, p)
        rt, rw = oracle_lib.bind_tets(V, elems, m.inv_dm, p)
        np.testing.assert_array_equal(t, rt)
        np.testing.assert_array_equal(w.view(np.uint32), rw.view(np.uint32))
    assert mesh.num_tets == 2


def test_reference_box_cage_binding(gpu):
    """mod.rs:523-560: weights sum to 1 and are >= -1e-4 inside the box; the
    rest position is reconstructed from the bound tet within 1e-4."""
    cage = _box([-1, -1, -1], [1, 1, 1])
    t, w = cage.bind_gaussians_gpu(gpu, np.array([[0, 0, 0], [0.5, -0.5, 0.5]], np.float32))
    assert np.all(np.abs(w.sum(1) - 1.0) < 1e-5) and np.all(w >= -1e-4)
    cage = _box([0, 0, 0], [2, 2, 2])
    pt = np.array([[0.7, 1.3, 0.4]], np.float32)
    t, w = cage.bind_gaussians_gpu(gpu, pt)
    rec = (w[0][:, None] * cage.rest_vertices[cage.elements[t[0]]]).sum(0)
    assert np.all(np.abs(rec - pt[0]) < 1e-4)


def test_reference_multi_cage_auto_bind(gpu):
    """mod.rs:562-600: first cage wins, disjoint, the far point stays unbound."""
    a = _box([-2.0, -1.0, -1.0], [-0.5, 1.0, 1.0])
    b = _box([0.5, -1.0, -1.0], [2.0, 1.0, 1.0])
    pos = np.array([[-1, 0, 0], [-0.8, 0.2, 0.1], [1, 0, 0], [0.9, -0.3, 0.2], [5, 5, 5]],
                   np.float32)
    cage, tet, w = auto_bind_gpu(gpu, [a, b], pos)
    assert cage.tolist() == [0, 0, 1, 1, UNBOUND]


def test_scene_binding_1p07m(gpu, oracle_lib):
    """The 1,073,930-splat scene on its 6000-tet cage: every splat bound, a
    sample of 40k checked against the oracle bit for bit."""
    from cindergs_amd import synthetic as syn

    sc = syn.scene_clusters()
    lo, hi = sc.positions.min(0), sc.positions.max(0)
    mesh = TetMesh.jittered_grid(lo, hi)
    t, w = mesh.bind_gaussians_gpu(gpu, sc.positions)
    assert np.all(t != UNBOUND)
    idx = np.random.default_rng(0).choice(sc.count, 40_000, replace=False)
    rt, rw = oracle_lib.bind_tets(mesh.rest_vertices, mesh.elements, mesh.inv_dm,
                                  sc.positions[idx])
    np.testing.assert_array_equal(t[idx], rt)
    np.testing.assert_array_equal(w[idx].view(np.uint32), rw.view(np.uint32))
