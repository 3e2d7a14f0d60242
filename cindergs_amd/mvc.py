"""Closed triangle cages and mean value coordinates on the GPU.

Host mirror of PolyCageMesh (src/deformation/mvc.rs): the cage data, the
signed volume / closed-manifold checks (mvc.rs:290-324) and calls into the
HIP library for the exact MVC binding (bind_gaussians, 498-541), the inside
test (winding_number, 326-363), the deformation of bound points
(evaluate_position + local_jacobian, 66-78 / 473-481) and of whole scenes
(deform_scene_cpu, 564-749).

The workload cage is a subdivided icosphere (642 vertices, 1280 triangles:
three 1-to-4 subdivisions of an icosahedron) of radius 1.2 with seeded radial
noise U(-0.15, 0.15), and the points are uniform inside it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import rng as R
from .cages import TetMesh
from .pipeline import DeviceBuffer, GpuPipelineState, _ptr, check

SEED_BASE = 0xC1DE5EED00000000


@dataclass
class PolyCage:
    vertices: np.ndarray   # [nv, 3] f32 rest positions
    triangles: np.ndarray  # [nt, 3] u32, outward (counter-clockwise seen from outside)

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(self.triangles, dtype=np.uint32).reshape(-1, 3)

    @property
    def num_vertices(self) -> int:
        return int(self.vertices.shape[0])

    def signed_volume(self) -> float:
        """Divergence theorem, positive for outward winding (mvc.rs:313-324)."""
        v = self.vertices.astype(np.float64)
        a, b, c = v[self.triangles[:, 0]], v[self.triangles[:, 1]], v[self.triangles[:, 2]]
        return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)

    def is_closed_manifold(self) -> bool:
        """Every half-edge (a, b) appears once and is paired with (b, a)."""
        he = {}
        for a, b, c in self.triangles.tolist():
            for e in ((a, b), (b, c), (c, a)):
                if e in he:
                    return False
                he[e] = 1
        return all((v, u) in he for (u, v) in he)

    def diagonal(self) -> float:
        return float(np.linalg.norm(self.vertices.max(0) - self.vertices.min(0)))

    def to_tet_mesh(self):
        """PolyCageMesh::to_tet_mesh (mvc.rs:751-779): the vertices plus their f32
        centroid (summed in vertex order, times 1/n), and one tet per triangle
        fanned to the centroid, (i0, i1, i2, c) when det[v1-v0, v2-v0, c-v0] >= 0,
        else (i0, i2, i1, c).  The determinant is glam's z . (x cross y) in f32."""
        f = np.float32
        V = self.vertices
        acc = np.zeros(3, f)
        for v in V:
            acc = (acc + v).astype(f)
        centroid = (acc * (f(1.0) / f(V.shape[0]))).astype(f)
        v0, v1, v2 = V[self.triangles[:, 0]], V[self.triangles[:, 1]], V[self.triangles[:, 2]]
        x, y, z = (v1 - v0).astype(f), (v2 - v0).astype(f), (centroid - v0).astype(f)
        cx = x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1]
        cy = x[:, 2] * y[:, 0] - x[:, 0] * y[:, 2]
        cz = x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0]
        det = (z[:, 0] * cx + z[:, 1] * cy) + z[:, 2] * cz
        c = np.full(self.triangles.shape[0], V.shape[0], np.uint32)
        t = self.triangles
        keep = det >= 0
        elems = np.stack([t[:, 0], np.where(keep, t[:, 1], t[:, 2]),
                          np.where(keep, t[:, 2], t[:, 1]), c], 1).astype(np.uint32)
        return TetMesh(np.concatenate([V, centroid[None]]).astype(f), elems)


def icosahedron(radius: float = 1.0) -> PolyCage:
    """Regular icosahedron: the 12 cyclic permutations of (0, +-1, +-phi) on a
    sphere, faces = triples of mutually adjacent vertices, wound outward."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    raw = []
    for s1 in (-1.0, 1.0):
        for s2 in (-1.0, 1.0):
            raw += [(0.0, s1, s2 * phi), (s1, s2 * phi, 0.0), (s2 * phi, 0.0, s1)]
    v = np.array(raw, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    d = np.linalg.norm(v[:, None] - v[None], axis=2)
    edge = np.sort(d[0])[1]
    adj = np.abs(d - edge) < 1e-6
    faces = []
    for i in range(12):
        for j in range(i + 1, 12):
            for k in range(j + 1, 12):
                if adj[i, j] and adj[j, k] and adj[i, k]:
                    n = np.cross(v[j] - v[i], v[k] - v[i])
                    faces.append((i, j, k) if np.dot(n, v[i] + v[j] + v[k]) > 0 else (i, k, j))
    return PolyCage((v * radius).astype(np.float32), np.array(faces, dtype=np.uint32))


def icosphere(subdivisions: int = 3, radius: float = 1.2, noise: float = 0.15,
              seed: int = 4) -> PolyCage:
    """Icosahedron subdivided 1-to-4 `subdivisions` times with midpoints pushed
    to the sphere (3 -> 642 vertices, 1280 triangles), then each vertex moved
    radially by U(-noise, noise) (seeded)."""
    base = icosahedron(1.0)
    v = [tuple(p) for p in base.vertices.astype(np.float64)]
    faces = [tuple(f) for f in base.triangles.tolist()]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = (np.array(v[a]) + np.array(v[b])) * 0.5
                p /= np.linalg.norm(p)
                mid[key] = len(v)
                v.append(tuple(p))
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(v, dtype=np.float64)
    # vertex v's radius: radius - noise + 2 noise u, u from stream (SEED_BASE +
    # seed, D_ICOSPHERE, v)
    u = R.Streams(SEED_BASE + seed, R.D_ICOSPHERE, V.shape[0]).uniform(1)[:, 0]
    r = radius - noise + 2.0 * noise * u
    return PolyCage((V * r[:, None]).astype(np.float32), np.array(faces, dtype=np.uint32))


class MvcGpu:
    """Exact MVC on one GpuPipelineState's device (cgs_mvc_* entry points)."""

    def __init__(self, state: GpuPipelineState, cage: PolyCage):
        self.state = state
        self.cage = cage
        self._lib = state._lib

    def winding(self, d_points: DeviceBuffer, n: int) -> np.ndarray:
        out = self.state.to_device(np.zeros(n, np.float32))
        check(self._lib.cgs_mvc_winding(self.state._h, _ptr(self.cage.vertices),
                                        self.cage.num_vertices, _ptr(self.cage.triangles),
                                        self.cage.triangles.shape[0], d_points.ptr, n, out.ptr))
        w = out.numpy()
        out.free()
        return w

    def bind(self, d_points: DeviceBuffer, n: int, gradients: bool = True):
        """bind_gaussians: device (weights [n, nv], gradients [n, nv, 3] or None,
        inside [n] u32) buffers."""
        nv = self.cage.num_vertices
        w = self.state.to_device(np.zeros((n, nv), np.float32)) if n else None
        g = self.state.to_device(np.zeros((n, nv, 3), np.float32)) if (n and gradients) else None
        m = self.state.to_device(np.zeros(n, np.uint32)) if n else None
        if n:
            check(self._lib.cgs_mvc_bind_exact(
                self.state._h, _ptr(self.cage.vertices), nv, _ptr(self.cage.triangles),
                self.cage.triangles.shape[0], d_points.ptr, n, w.ptr, g.ptr if g else None,
                m.ptr))
        return w, g, m

    def far_share(self, d_points: DeviceBuffer, n: int, r0: float, r1: float,
                  kind: int = 0) -> DeviceBuffer:
        """Near/far training targets (cgs_mvc_far_targets): [n, nv] device rows
        of nv * (u - u_near) / sum u (kind 0, CGS_FAR_SHARE) or of
        (u - u_near) * nv * diagonal / (4 pi) (kind 1, CGS_FAR_SUMS)."""
        nv = self.cage.num_vertices
        out = self.state.to_device(np.zeros((n, nv), np.float32))
        if n:
            check(self._lib.cgs_mvc_far_targets(self.state._h, _ptr(self.cage.vertices), nv,
                                                _ptr(self.cage.triangles),
                                                self.cage.triangles.shape[0], float(r0), float(r1),
                                                d_points.ptr, n, int(kind), out.ptr))
        return out

    def nearfar_bind(self, d_points: DeviceBuffer, n: int, d_q: DeviceBuffer, r0: float, r1: float,
                     d_dq=None, gradients: bool = True, kind: int = 0):
        """The near/far combination with given far targets of `kind`
        (cgs_mvc_nearfar_bind_kind): device (weights, gradients or None,
        inside, residual)."""
        nv = self.cage.num_vertices
        w = self.state.to_device(np.zeros((n, nv), np.float32))
        g = self.state.to_device(np.zeros((n, nv, 3), np.float32)) if gradients else None
        m = self.state.to_device(np.zeros(n, np.uint32))
        r = self.state.to_device(np.zeros(n, np.float32))
        check(self._lib.cgs_mvc_nearfar_bind_kind(
            self.state._h, _ptr(self.cage.vertices), nv, _ptr(self.cage.triangles),
            self.cage.triangles.shape[0], float(r0), float(r1), d_points.ptr, n, d_q.ptr,
            d_dq.ptr if d_dq is not None else None, int(kind), w.ptr, g.ptr if g else None, m.ptr,
            r.ptr))
        return w, g, m, r

    def deform(self, w: DeviceBuffer, g, m: DeviceBuffer, d_rest: DeviceBuffer,
               deformed_vertices: np.ndarray, n: int, jacobians: bool = True):
        dv = np.ascontiguousarray(deformed_vertices, dtype=np.float32)
        pos = self.state.to_device(np.zeros((n, 3), np.float32))
        F = self.state.to_device(np.zeros((n, 9), np.float32)) if jacobians else None
        check(self._lib.cgs_mvc_deform(self.state._h, w.ptr, g.ptr if g else None,
                                       m.ptr if m is not None else None, d_rest.ptr, _ptr(dv), self.cage.num_vertices, n, pos.ptr,
                                       F.ptr if F else None))
        return pos, F

    def deform_scene(self, w: DeviceBuffer, g: DeviceBuffer, m, rest: dict,
                     deformed_vertices: np.ndarray, n: int, sh_degree: int = 3,
                     split_threshold: float = 0.0, mip_filter_3d_var: float = 0.0,
                     out: dict | None = None):
        """PolyCageMesh::deform_scene_cpu (mvc.rs:564-749) on the device.

        rest: device buffers "pos" [n,3], "cov" [n,6], "opa" [n], "sh" [n,48];
        m: the inside flags of the binding (or None: every splat bound).  Returns
        (count, out) where out holds device buffers "pos", "cov", "opa", "sh"
        of 2n records (preallocated ones may be passed in), the first `count`
        valid: the n deformed records, then the second child of each split
        splat in parent order."""
        dv = np.ascontiguousarray(deformed_vertices, dtype=np.float32)
        if out is None:
            cap = max(2 * n, 1)
            out = {"pos": self.state.to_device(np.zeros((cap, 3), np.float32)),
                   "cov": self.state.to_device(np.zeros((cap, 6), np.float32)),
                   "opa": self.state.to_device(np.zeros(cap, np.float32)),
                   "sh": self.state.to_device(np.zeros((cap, 48), np.float32))}
        count = C.c_uint64(0)
        check(self._lib.cgs_mvc_deform_scene(
            self.state._h, w.ptr if w is not None else None, g.ptr if g is not None else None,
            m.ptr if m is not None else None, _ptr(dv),
            self.cage.num_vertices, n, rest["pos"].ptr, rest["cov"].ptr, rest["opa"].ptr,
            rest["sh"].ptr, int(sh_degree), float(split_threshold), float(mip_filter_3d_var),
            out["pos"].ptr, out["cov"].ptr, out["opa"].ptr, out["sh"].ptr, C.byref(count)))
        return int(count.value), out


def sample_inside(cage: PolyCage, n: int, seed: int = 4, inside_fn=None,
                  batch: int = 1 << 20) -> np.ndarray:
    """n points uniform inside the cage: rejection from the bounding box by
    |winding number| >= 0.5 (inside_fn(points) -> winding numbers).  Candidate
    c is lo + (hi - lo) u (cast to f32), u the 3 uniforms of stream
    (SEED_BASE + seed + 17, D_SAMPLE, c); candidates are taken in order."""
    lo, hi = cage.vertices.min(0).astype(np.float64), cage.vertices.max(0).astype(np.float64)
    out = []
    have = 0
    first = 0
    while have < n:
        cand = R.uniform(SEED_BASE + seed + 17, R.D_SAMPLE, batch, 3, lo, hi, first).astype(np.float32)
        first += batch
        keep = cand[np.abs(inside_fn(cand)) >= 0.5]
        out.append(keep)
        have += keep.shape[0]
    return np.ascontiguousarray(np.concatenate(out)[:n])
