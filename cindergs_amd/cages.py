"""Volumetric tetrahedral cages on the host: construction, splat binding and the
spring simulator that drives the cage vertices frame by frame.

Follows src/deformation/mod.rs (TetMesh::new 202-227: D_m = [X1-X0, X2-X0,
X3-X0], inverse stored column-major, identity when |det| <= 1e-8;
bind_point 267-310: the first tetrahedron in index order whose barycentric
penalty sum(max(-w_k, 0)) is <= 1e-4, else the one with the smallest penalty)
and src/deformation/dynamics.rs (CageSpringSimulator, 74-316: semi-implicit
Euler, a = k (x_rest - x) / m, v = (v + a dt) * exp(-c dt), pins and
collision planes).

The grid cage is a jittered Freudenthal (Kuhn) subdivision: each cell of an
nx x ny x nz grid over the scene's bounds is split into 6 tetrahedra sharing
the cell diagonal; 10 x 10 x 10 cells give 6,000 tetrahedra.
"""
from __future__ import annotations

import itertools
import math

import numpy as np

from . import rng
from .pipeline import TetCage

UNBOUND = 0xFFFFFFFF

# the 6 tetrahedra of a Kuhn-subdivided cube (corner index bit0=x, bit1=y, bit2=z),
# all containing the diagonal 0 -> 7 and positively oriented
_KUHN = []
for perm in itertools.permutations((1, 2, 4)):
    a = perm[0]
    b = a | perm[1]
    _KUHN.append((0, a, b, 7))


def _cross(a, b):
    """glam Vec3::cross, f32, the crate's operation order."""
    return np.stack([a[:, 1] * b[:, 2] - b[:, 1] * a[:, 2], a[:, 2] * b[:, 0] - b[:, 2] * a[:, 0],
                     a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1]], 1)


def _dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def inverse_rest_matrices(rest_vertices: np.ndarray, elements: np.ndarray) -> np.ndarray:
    """D_m^-1 per tet as TetMesh::new computes it (src/deformation/mod.rs:202-227):
    D_m = [X1-X0, X2-X0, X3-X0] in f32, glam 0.28's Mat3::inverse (column cross
    products, det = z . (x cross y), times 1/det, transposed) when |det| > 1e-8,
    else the identity; [T, 9] column-major (Mat3::to_cols_array)."""
    X = np.asarray(rest_vertices, dtype=np.float32)
    e = np.asarray(elements, dtype=np.int64)
    x0 = X[e[:, 0]]
    cx, cy, cz = X[e[:, 1]] - x0, X[e[:, 2]] - x0, X[e[:, 3]] - x0
    t0, t1, t2 = _cross(cy, cz), _cross(cz, cx), _cross(cx, cy)
    det = _dot(cz, t2)
    ok = np.abs(det) > np.float32(1e-8)
    inv = (np.float32(1.0) / np.where(ok, det, np.float32(1.0))).astype(np.float32)
    r0, r1, r2 = t0 * inv[:, None], t1 * inv[:, None], t2 * inv[:, None]
    # transposed: column j of the inverse holds component j of r0, r1, r2
    out = np.stack([r0[:, 0], r1[:, 0], r2[:, 0], r0[:, 1], r1[:, 1], r2[:, 1],
                    r0[:, 2], r1[:, 2], r2[:, 2]], 1).astype(np.float32)
    out[~ok] = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    return np.ascontiguousarray(out)


class TetMesh:
    """TetMesh (src/deformation/mod.rs:191-227) with a cell index for binding."""

    def __init__(self, rest_vertices: np.ndarray, elements: np.ndarray, cells=None, lo=None,
                 hi=None):
        self.rest_vertices = np.ascontiguousarray(rest_vertices, dtype=np.float32)
        self.elements = np.ascontiguousarray(elements, dtype=np.uint32)
        self.cells = cells
        self.lo = lo
        self.hi = hi
        self.inv_dm = inverse_rest_matrices(self.rest_vertices, self.elements)

    @property
    def num_tets(self) -> int:
        return int(self.elements.shape[0])

    @property
    def num_vertices(self) -> int:
        return int(self.rest_vertices.shape[0])

    @staticmethod
    def jittered_grid(lo, hi, cells=(10, 10, 10), jitter: float = 0.15, seed: int = 3,
                      pad: float = 1e-3) -> "TetMesh":
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        ext = np.maximum(hi - lo, 1e-3)
        lo = lo - ext * pad
        hi = hi + ext * pad
        nx, ny, nz = cells
        gx = np.linspace(lo[0], hi[0], nx + 1)
        gy = np.linspace(lo[1], hi[1], ny + 1)
        gz = np.linspace(lo[2], hi[2], nz + 1)
        V = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), -1).reshape(-1, 3)
        h = (hi - lo) / np.array(cells)
        # vertex v draws 3 uniforms from stream (0xC1DE5EED00000000 + seed,
        # D_GRID, v); interior vertices move by (-jitter + 2 jitter u) h
        ijk = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                                   indexing="ij"), -1).reshape(-1, 3)
        interior = np.all((ijk > 0) & (ijk < np.array(cells)), axis=1)
        u = rng.Streams(0xC1DE5EED00000000 + seed, rng.D_GRID, V.shape[0]).uniform(3)
        V[interior] += (-jitter + 2.0 * jitter * u[interior]) * h

        def vid(i, j, k):
            return (i * (ny + 1) + j) * (nz + 1) + k

        elems = []
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    corner = [vid(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))
                              for c in range(8)]
                    for t in _KUHN:
                        elems.append([corner[t[0]], corner[t[1]], corner[t[2]], corner[t[3]]])
        return TetMesh(V.astype(np.float32), np.array(elems, dtype=np.uint32), cells, lo, hi)

    def barycentric(self, p: np.ndarray, tets: np.ndarray) -> np.ndarray:
        """[w0..w3] of points p[n,3] in tets[n] with bind_point's f32 operation
        order (mod.rs:280-290): w_r = (m[0,r] dx + m[1,r] dy) + m[2,r] dz."""
        X0 = self.rest_vertices[self.elements[tets, 0]]
        m = self.inv_dm[tets]  # column-major: m[:, col*3 + row]
        d = (p.astype(np.float32) - X0).astype(np.float32)
        w = [(m[:, r] * d[:, 0] + m[:, 3 + r] * d[:, 1]) + m[:, 6 + r] * d[:, 2]
             for r in range(3)]
        w0 = np.float32(1.0) - ((w[0] + w[1]) + w[2])
        return np.stack([w0, w[0], w[1], w[2]], 1).astype(np.float32)

    def bounds(self):
        """TetMesh::bounds (mod.rs:242-256): AABB padded by 1e-3 of its size."""
        lo = self.rest_vertices.min(0)
        hi = self.rest_vertices.max(0)
        pad = np.maximum(hi - lo, np.float32(1e-3)) * np.float32(1e-3)
        return (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)

    def bind_gaussians_gpu(self, state, positions: np.ndarray, bounds=None):
        """bind_gaussians(_with_bounds) on the GPU (grid-accelerated, same result
        as the reference's index-order scan)."""
        return state.bind_tets(self.rest_vertices, self.elements, self.inv_dm, positions, bounds)

    def bind_gaussians(self, positions: np.ndarray):
        """bind_point for every position: first tet in index order with penalty
        <= 1e-4, else minimal penalty.  On a grid cage only the tetrahedra of
        the 27 cells around the point's cell can contain it (jitter < half a
        cell), so those are the candidates searched, in index order; any other
        mesh is scanned whole."""
        pos = np.asarray(positions, dtype=np.float32)
        if self.cells is None:
            return self._bind_scan(pos)
        n = pos.shape[0]
        nx, ny, nz = self.cells
        cell = np.floor((pos - self.lo) / ((self.hi - self.lo) / np.array(self.cells))).astype(np.int64)
        cell = np.clip(cell, 0, np.array(self.cells) - 1)
        best_t = np.zeros(n, dtype=np.int64)
        best_pen = np.full(n, np.inf, dtype=np.float32)
        best_w = np.zeros((n, 4), dtype=np.float32)
        first_t = np.full(n, np.iinfo(np.int64).max, dtype=np.int64)
        first_w = np.zeros((n, 4), dtype=np.float32)
        for di, dj, dk in itertools.product((-1, 0, 1), repeat=3):
            c = cell + np.array([di, dj, dk])
            ok = np.all((c >= 0) & (c < np.array(self.cells)), axis=1)
            cidx = (c[:, 0] * ny + c[:, 1]) * nz + c[:, 2]
            for s in range(6):
                t = np.where(ok, cidx * 6 + s, 0)
                w = self.barycentric(pos, t)
                pen = (np.maximum(-w, 0).sum(1)).astype(np.float32)
                hit = ok & (pen <= 1e-4) & (t < first_t)
                first_t = np.where(hit, t, first_t)
                first_w[hit] = w[hit]
                better = ok & ((pen < best_pen) | ((pen == best_pen) & (t < best_t)))
                best_pen = np.where(better, pen, best_pen)
                best_t = np.where(better, t, best_t)
                best_w[better] = w[better]
        found = first_t != np.iinfo(np.int64).max
        tet = np.where(found, first_t, best_t).astype(np.uint32)
        w = np.where(found[:, None], first_w, best_w).astype(np.float32)
        return tet, np.ascontiguousarray(w)

    def _bind_scan(self, pos: np.ndarray):
        """bind_point over every tet in index order (mod.rs:267-305)."""
        n = pos.shape[0]
        best_t = np.zeros(n, np.int64)
        best_pen = np.full(n, np.inf, np.float32)
        best_w = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
        found = np.zeros(n, bool)
        first_t = np.zeros(n, np.int64)
        first_w = np.zeros((n, 4), np.float32)
        for t in range(self.num_tets):
            w = self.barycentric(pos, np.full(n, t, np.int64))
            pen = np.maximum(-w, 0).sum(1).astype(np.float32)
            hit = ~found & (pen <= 1e-4)
            first_t[hit] = t
            first_w[hit] = w[hit]
            found |= hit
            better = pen < best_pen
            best_pen = np.where(better, pen, best_pen)
            best_t = np.where(better, t, best_t)
            best_w[better] = w[better]
        tet = np.where(found, first_t, best_t).astype(np.uint32)
        return tet, np.ascontiguousarray(np.where(found[:, None], first_w, best_w), np.float32)

    def cage(self, vertices: np.ndarray, bind_tet: np.ndarray, bind_w: np.ndarray) -> TetCage:
        return TetCage(self.elements, self.inv_dm, np.ascontiguousarray(vertices, dtype=np.float32),
                       np.ascontiguousarray(bind_tet, dtype=np.uint32),
                       np.ascontiguousarray(bind_w, dtype=np.float32))


F32 = np.float32


def _dot3(a, b):
    """glam Vec3::dot: (x*x' + y*y') + z*z' in f32"""
    return F32(F32(F32(a[0] * b[0]) + F32(a[1] * b[1])) + F32(a[2] * b[2]))


class CollisionPlane:
    """CollisionPlane (src/deformation/dynamics.rs:3-72): the half-space
    n . x >= offset, with restitution and tangential friction on contact."""

    def __init__(self, normal, offset: float, restitution: float, friction: float):
        n = np.asarray(normal, F32).reshape(3)
        ls = _dot3(n, n)
        if ls > F32(1e-8):
            n = (n * (F32(1.0) / np.sqrt(ls, dtype=F32))).astype(F32)  # Vec3::normalize
        else:
            n = np.array([0, 1, 0], F32)
        self.normal = n
        self.offset = F32(offset)
        self.restitution = F32(min(max(restitution, 0.0), 1.0))
        self.friction = F32(min(max(friction, 0.0), 1.0))

    @classmethod
    def ground_y(cls, y: float) -> "CollisionPlane":
        return cls((0.0, 1.0, 0.0), y, 0.25, 0.2)

    def signed_distance(self, point) -> np.float32:
        return F32(_dot3(self.normal, np.asarray(point, F32)) - self.offset)

    def clamp_point(self, point) -> np.ndarray:
        p = np.asarray(point, F32)
        d = self.signed_distance(p)
        return (p - self.normal * d).astype(F32) if d < 0 else p

    def resolve_contact(self, pos: np.ndarray, vel: np.ndarray) -> bool:
        """resolve_contact (47-60) on one vertex, in place."""
        d = self.signed_distance(pos)
        if d >= 0:
            return False
        pos -= self.normal * d
        vn = _dot3(vel, self.normal)
        if vn < 0:
            v_norm = (self.normal * vn).astype(F32)
            v_tan = (vel - v_norm).astype(F32)
            vel[:] = (v_tan * (F32(1.0) - self.friction) - v_norm * self.restitution).astype(F32)
        return True


class CageSpringSimulator:
    """CageSpringSimulator (src/deformation/dynamics.rs:74-316), in f32.

    Hookean springs towards the rest cage, semi-implicit Euler with the
    frame-rate independent damping gamma = exp(-c dt), c = -ln(retention) * 60
    (`damping` is the velocity retention per 1/60 s frame), dt clamped to
    MAX_DT, transient (`pinned_index`) and persistent pins, and collision
    planes.  Defaults are the reference's (stiffness 180, retention 0.92, mass
    1).  The cage workload's spring (stiffness 40, c = 4 /s) is
    `CageSpringSimulator(rest, stiffness=40.0, damping=math.exp(-4.0 / 60.0))`.
    """

    MAX_DT = F32(0.033)
    DAMPING_REFERENCE_HZ = F32(60.0)

    def __init__(self, rest: np.ndarray, stiffness: float = 180.0, damping: float = 0.92,
                 mass: float = 1.0):
        self.rest_positions = np.ascontiguousarray(rest, dtype=F32).reshape(-1, 3).copy()
        self.current_positions = self.rest_positions.copy()
        self.velocities = np.zeros_like(self.rest_positions)
        self.stiffness = F32(stiffness)
        self.damping = F32(damping)
        self.mass = F32(mass)
        self.pinned_index = None
        self.pinned_vertices = np.zeros(self.rest_positions.shape[0], bool)
        self.collision_planes: list[CollisionPlane] = []

    # -- the names the renderer's callers use ---------------------------------
    @property
    def x(self) -> np.ndarray:
        """a copy of the current positions (step updates them in place)"""
        return self.current_positions.copy()

    @property
    def rest(self) -> np.ndarray:
        return self.rest_positions

    @staticmethod
    def damping_coefficient(damping: float) -> np.float32:
        """-ln(clamp(retention, 1e-4, 0.9999)) * 60 (138-141)"""
        r = F32(min(max(F32(damping), F32(1.0e-4)), F32(0.9999)))
        return F32(-np.log(r, dtype=F32) * CageSpringSimulator.DAMPING_REFERENCE_HZ)

    @staticmethod
    def frame_damping(damping: float, dt: float) -> np.float32:
        """exp(-c dt) (145-148)"""
        c = CageSpringSimulator.damping_coefficient(damping)
        return F32(np.exp(F32(-c * F32(dt)), dtype=F32))

    def is_vertex_pinned(self, i: int) -> bool:
        return self.pinned_index == i or bool(self.pinned_vertices[i])

    def pin_vertex(self, i: int) -> None:
        if i < self.pinned_vertices.size:
            self.pinned_vertices[i] = True
            self.velocities[i] = 0

    def unpin_vertex(self, i: int) -> None:
        if i < self.pinned_vertices.size:
            self.pinned_vertices[i] = False

    def toggle_pin_vertex(self, i: int) -> bool:
        if i >= self.pinned_vertices.size:
            return False
        self.pinned_vertices[i] = not self.pinned_vertices[i]
        if self.pinned_vertices[i]:
            self.velocities[i] = 0
        return bool(self.pinned_vertices[i])

    def clear_pins(self) -> None:
        self.pinned_vertices[:] = False

    def pinned_count(self) -> int:
        m = self.pinned_vertices.copy()
        if self.pinned_index is not None and self.pinned_index < m.size:
            m[self.pinned_index] = True
        return int(m.sum())

    def add_collision_plane(self, plane: CollisionPlane) -> None:
        self.collision_planes.append(plane)

    def clear_collision_planes(self) -> None:
        self.collision_planes.clear()

    def project_point_to_collision_planes(self, point) -> np.ndarray:
        p = np.asarray(point, F32)
        for pl in self.collision_planes:
            p = pl.clamp_point(p)
        return p

    def set_pinned(self, pinned) -> None:
        self.pinned_index = pinned
        if pinned is not None and pinned < self.velocities.shape[0]:
            self.velocities[pinned] = 0

    def set_pinned_position(self, i: int, pos) -> None:
        if i < self.current_positions.shape[0]:
            self.current_positions[i] = self.project_point_to_collision_planes(pos)
            self.velocities[i] = 0

    def step(self, dt: float) -> np.ndarray:
        """step (249-277): the per-vertex update of the reference, vectorised
        over the vertices with the same f32 operations in the same order."""
        dt_sim = F32(min(F32(dt), self.MAX_DT))
        if dt_sim <= 0:
            return self.current_positions
        inv_mass = F32(F32(1.0) / max(self.mass, F32(0.001)))
        k = self.stiffness
        gamma = self.frame_damping(self.damping, dt_sim)
        free = ~self.pinned_vertices.copy()
        if self.pinned_index is not None and self.pinned_index < free.size:
            free[self.pinned_index] = False
        self.velocities[~free] = 0
        x, v = self.current_positions, self.velocities
        disp = self.rest_positions[free] - x[free]
        accel = disp * F32(k * inv_mass)
        v[free] = (v[free] + accel * dt_sim) * gamma
        x[free] = x[free] + v[free] * dt_sim
        if self.collision_planes:
            for i in np.nonzero(free)[0]:
                for pl in self.collision_planes:
                    pl.resolve_contact(x[i], v[i])
        return x

    def trigger_jiggle_impulse(self, magnitude: float) -> None:
        """Outward radial velocity pulse on every unpinned vertex (280-306)."""
        n = self.rest_positions.shape[0]
        if n == 0:
            return
        c = np.zeros(3, F32)
        for p in self.rest_positions:
            c = (c + p).astype(F32)
        c = (c / F32(n)).astype(F32)
        mag = F32(magnitude)
        for i in range(n):
            if self.is_vertex_pinned(i):
                continue
            d = (self.rest_positions[i] - c).astype(F32)
            ln = np.sqrt(_dot3(d, d), dtype=F32)
            rcp = F32(1.0) / ln if ln > 0 else F32(np.inf)
            d = (d * rcp).astype(F32) if np.isfinite(rcp) and rcp > 0 else np.zeros(3, F32)
            imp = d * mag if _dot3(d, d) > F32(1e-4) else np.array([0, mag, 0], F32)
            self.velocities[i] = (self.velocities[i] + imp).astype(F32)

    def reset(self) -> None:
        self.current_positions[:] = self.rest_positions
        self.velocities[:] = 0
        self.pinned_index = None

    def positions_array(self) -> np.ndarray:
        return self.current_positions.copy()

    def impulse(self, magnitude: float = 0.2, fraction: float = 0.05, seed: int = 3) -> None:
        """The cage workload's excitation (no reference counterpart): an outward
        radial velocity pulse of `magnitude` on a seeded `fraction` of the
        vertices, directions as in trigger_jiggle_impulse.  Vertex v draws one
        uniform from stream (0xC1DE5EED00000000 + seed + 100, D_IMPULSE, v);
        the vertices with the smallest draws are picked."""
        n = self.rest_positions.shape[0]
        u = rng.Streams(0xC1DE5EED00000000 + seed + 100, rng.D_IMPULSE, n).uniform(1)[:, 0]
        pick = np.sort(np.argsort(u, kind="stable")[:max(1, int(round(fraction * n)))])
        centroid = self.rest_positions.mean(0)
        d = self.rest_positions[pick] - centroid
        ln = np.linalg.norm(d, axis=1, keepdims=True)
        d = np.where(ln > 1e-2, d / np.maximum(ln, 1e-12), np.array([0, 1, 0], dtype=F32))
        self.velocities[pick] += (d * magnitude).astype(F32)


def auto_bind_gpu(state, meshes, positions: np.ndarray):
    """CageScene::auto_bind (src/deformation/mod.rs:444-464): each cage binds the
    splats inside its padded rest bounds; the first cage in the list keeps a
    splat bound by several.  Returns (cage[n] u32, tet[n] u32, weights[n,4]);
    cage = 0xFFFFFFFF for an unbound splat."""
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    n = p.shape[0]
    cage = np.full(n, UNBOUND, np.uint32)
    tet = np.zeros(n, np.uint32)
    w = np.zeros((n, 4), np.float32)
    for ci, m in enumerate(meshes):
        t, ww = m.bind_gaussians_gpu(state, p, m.bounds())
        take = (cage == UNBOUND) & (t != UNBOUND)
        cage[take] = ci
        tet[take] = t[take]
        w[take] = ww[take]
    return cage, tet, w
