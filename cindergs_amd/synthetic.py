"""Seeded synthetic scenes and cameras for the benchmark workloads.

The reference has no generator for these distributions; the specification
here is ours:
  * RNG: the streams of cindergs_amd.rng (SplitMix64-seeded xoshiro256**,
    one stream per (seed, domain, index), Box-Muller normals in float64) with
    seed 0xC1DE5EED0000000k for workload k.
  * Splats: log-scales ~ N(-4.5, 0.5); unit quaternions (PLY WXYZ order,
    src/math/covariance.rs:36-51) from a normalised N(0, I4); opacity =
    sigmoid(N(0, 1.5)); SH degree 3 (48 f32, src/render/rasterizer/sh.rs:56-61)
    with DC ~ N(0, 0.5) and higher orders ~ N(0, 0.05); covariance
    Sigma = (R S)(R S)^T (covariance.rs:60-65) computed in f64, stored f32.
  * Means: N(0, 0.25 I) for the 100k scene; a mixture of 32 anisotropic
    Gaussian clusters (centres U[-3,3]^3, axis radii U(0.05, 0.6), random
    orientation) plus 10 % uniform background in [-4,4]^3 for the 1.07M scene.
  * Cameras: glam look_at_lh (camera.rs:71-96), square pixels,
    f = H / (2 tan(fov/2)), principal point at the image centre, near 0.2.
The draw order of every stream is stated in each generator's docstring.
"""
from __future__ import annotations

import math

import numpy as np

from . import rng as R
from .pipeline import (CameraUniforms, CompositorUniforms, GaussianSceneSoa,
                       QuantizedGaussianSceneSoa)

SEED_BASE = 0xC1DE5EED00000000
CONFIG2_SPLATS = 1_073_930
ATTR_NORMALS = 56  # normals per splat for its attributes (see _attributes)


def quat_to_mat(q: np.ndarray) -> np.ndarray:
    """q [n,4] WXYZ unit quaternions -> [n,3,3] rotation matrices."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _attributes(z: np.ndarray, means: np.ndarray) -> GaussianSceneSoa:
    """A splat's attributes from its 56 standard normals z: log-scales
    -4.5 + 0.5 z[0:3], quaternion (WXYZ) normalise(z[3:7]), opacity
    sigmoid(1.5 z[7]), SH DC 0.5 z[8:11], SH rest 0.05 z[11:56]."""
    n = z.shape[0]
    s = np.maximum(np.exp(-4.5 + 0.5 * z[:, 0:3]), 1e-4)
    q = z[:, 3:7] / np.maximum(np.linalg.norm(z[:, 3:7], axis=1, keepdims=True), 1e-12)
    Rm = quat_to_mat(q)
    M = Rm * s[:, None, :]
    S = M @ np.transpose(M, (0, 2, 1))
    cov = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)
    opa = 1.0 / (1.0 + np.exp(-1.5 * z[:, 7]))
    sh = np.empty((n, 48))
    sh[:, :3] = 0.5 * z[:, 8:11]
    sh[:, 3:] = 0.05 * z[:, 11:56]
    return GaussianSceneSoa(np.ascontiguousarray(means.astype(np.float32)),
                            np.ascontiguousarray(cov.astype(np.float32)),
                            np.ascontiguousarray(opa.astype(np.float32)),
                            np.ascontiguousarray(sh.astype(np.float32)))


def _concat(parts) -> GaussianSceneSoa:
    return GaussianSceneSoa(*(np.ascontiguousarray(np.concatenate([getattr(p, f) for p in parts]))
                              for f in ("positions", "covariances_3d", "opacities", "sh_coeffs")))


def _splat_block(seed: int, first: int, n: int, fn) -> GaussianSceneSoa:
    """fn(streams) -> (means, z56) on splats first..first+n-1, in chunks."""
    from concurrent.futures import ThreadPoolExecutor
    import os

    chunk = 1 << 16
    starts = list(range(0, n, chunk))

    def one(c0):
        m, z = fn(R.Streams(seed, R.D_SPLAT, min(chunk, n - c0), first + c0))
        return _attributes(z, m)

    if len(starts) <= 1:
        return one(0) if n else _attributes(np.zeros((0, 56)), np.zeros((0, 3)))
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1, len(starts)))) as ex:
        return _concat(list(ex.map(one, starts)))


def scene_gaussian_ball(n: int = 100_000, k: int = 1) -> GaussianSceneSoa:
    """100k-splat scene: means ~ N(0, 0.5^2 I3).  Splat i draws 59 normals
    from stream (seed, D_SPLAT, i): mean 0.5 z[0:3], then its attributes
    from z[3:59]."""
    def fn(st):
        z = st.normal(59)
        return 0.5 * z[:, 0:3], z[:, 3:59]

    return _splat_block(SEED_BASE + k, 0, n, fn)


def scene_clusters(n: int = CONFIG2_SPLATS, k: int = 2, clusters: int = 32,
                   background: float = 0.10) -> GaussianSceneSoa:
    """Cluster-mixture scene (1,073,930 splats by default).  Cluster c draws
    from stream (seed, D_SCENE_GLOBAL, c) 6 uniforms (centre -3 + 6 u[0:3],
    axis radii 0.05 + 0.55 u[3:6]) and then 4 normals (orientation
    quaternion).  The first n - round(n background) splats are cluster
    members: splat i draws from (seed, D_SPLAT, i) one uniform (cluster
    floor(32 u)) and 59 normals (local offset z[0:3] times the radii,
    rotated, plus the centre; attributes from z[3:59]).  The rest are
    background: 3 uniforms (position -4 + 8 u) and 56 normals (attributes)."""
    seed = SEED_BASE + k
    gs = R.Streams(seed, R.D_SCENE_GLOBAL, clusters)
    u = gs.uniform(6)
    centres = -3.0 + 6.0 * u[:, 0:3]
    radii = 0.05 + 0.55 * u[:, 3:6]
    qc = gs.normal(4)
    qc /= np.linalg.norm(qc, axis=1, keepdims=True)
    Rc = quat_to_mat(qc)
    n_bg = int(round(n * background))
    n_cl = n - n_bg

    def member(st):
        which = np.minimum((st.uniform(1)[:, 0] * clusters).astype(np.int64), clusters - 1)
        z = st.normal(59)
        local = z[:, 0:3] * radii[which]
        pts = np.einsum("nij,nj->ni", Rc[which], local) + centres[which]
        return pts, z[:, 3:59]

    def back(st):
        pos = -4.0 + 8.0 * st.uniform(3)
        return pos, st.normal(56)

    return _concat([_splat_block(seed, 0, n_cl, member), _splat_block(seed, n_cl, n_bg, back)])


def look_at_lh(eye, target, up) -> np.ndarray:
    """glam Mat4::look_at_lh as a column-major 16-float array."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(np.asarray(up, dtype=np.float64), f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    return np.array([s[0], u[0], f[0], 0.0, s[1], u[1], f[1], 0.0, s[2], u[2], f[2], 0.0,
                     -eye.dot(s), -eye.dot(u), -eye.dot(f), 1.0])


def camera_look_at(eye, target, up, fov_y_deg: float, width: int, height: int,
                   near: float = 0.2) -> CameraUniforms:
    """CameraUniforms::from_look_at (camera.rs:71-96)."""
    cam = CameraUniforms()
    v = look_at_lh(eye, target, up)
    for i in range(16):
        cam.view_matrix[i] = float(np.float32(v[i]))
    f = 0.5 * height / math.tan(0.5 * math.radians(fov_y_deg))
    cam.focal_x = cam.focal_y = f
    cam.principal_x = 0.5 * width
    cam.principal_y = 0.5 * height
    cam.viewport_width = float(width)
    cam.viewport_height = float(height)
    cam.near_plane = near
    cam.pad = 0.0
    return cam


def compositor(width: int, height: int, eye, sh_degree: int = 3,
               background=(0.05, 0.05, 0.07, 1.0)) -> CompositorUniforms:
    """CompositorUniforms::new (rasterizer/mod.rs:44-60)."""
    c = CompositorUniforms()
    c.screen_width = width
    c.screen_height = height
    c.tiles_x = (width + 15) // 16
    c.tiles_y = (height + 15) // 16
    for i in range(4):
        c.background_color[i] = background[i]
    for i in range(3):
        c.camera_position[i] = float(eye[i])
    c.sh_degree = min(sh_degree, 3)
    return c


def orbit_eye(frame: int, frames: int = 120, radius: float = 8.0, height: float = 0.5):
    th = 2.0 * math.pi * frame / frames
    return (radius * math.sin(th), height, -radius * math.cos(th))


def view_config1():
    """One 800x600 camera, 60 deg fov, 3 units from the origin looking at it."""
    eye = (0.0, 0.0, -3.0)
    return camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, 800, 600), compositor(800, 600, eye)


def view_orbit(frame: int, width: int = 1920, height: int = 1080, frames: int = 120):
    """Frame `frame` of the 120-frame orbit (radius 8, height 0.5, fov 60 deg)."""
    eye = orbit_eye(frame, frames)
    return (camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, width, height),
            compositor(width, height, eye))


def camera_array(cam: CameraUniforms) -> np.ndarray:
    return np.frombuffer(bytes(cam), dtype=np.float32).copy()


def pack_f16_pair(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """soa.rs:348-352: low half = a, high half = b, round-to-nearest-even."""
    ha = a.astype(np.float16).view(np.uint16).astype(np.uint32)
    hb = b.astype(np.float16).view(np.uint16).astype(np.uint32)
    return ha | (hb << 16)


def quantize_f16(scene: GaussianSceneSoa, sh_degree: int = 3) -> QuantizedGaussianSceneSoa:
    """QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:454-500)."""
    deg = min(sh_degree, 3)
    bases = (deg + 1) ** 2
    words = (bases * 3 + 1) // 2
    n = scene.count
    cov = scene.covariances_3d
    cov_p = np.stack([pack_f16_pair(cov[:, 0], cov[:, 1]), pack_f16_pair(cov[:, 2], cov[:, 3]),
                      pack_f16_pair(cov[:, 4], cov[:, 5])], 1).astype(np.uint32)
    rest = bases - 1
    tmp = np.zeros((n, words * 2), dtype=np.float32)
    sh = scene.sh_coeffs
    tmp[:, 0:3] = sh[:, 0:3]
    if rest:
        tmp[:, 3:3 + rest] = sh[:, 3:3 + rest]
        tmp[:, 3 + rest:3 + 2 * rest] = sh[:, 18:18 + rest]
        tmp[:, 3 + 2 * rest:3 + 3 * rest] = sh[:, 33:33 + rest]
    sh_p = pack_f16_pair(tmp[:, 0::2], tmp[:, 1::2]).astype(np.uint32)
    return QuantizedGaussianSceneSoa(scene.positions.copy(), np.ascontiguousarray(cov_p),
                                     scene.opacities.copy(), np.ascontiguousarray(sh_p), deg)


BATCH_RESOLUTIONS = ((320, 240), (800, 600), (1920, 1080), (3840, 2160))
BATCH_SCENE_SPLATS = (200_000, CONFIG2_SPLATS, 4_000_000)


def batch_views(n_views: int = 64, n_scenes: int = 3, seed: int = 5,
                resolutions=BATCH_RESOLUTIONS):
    """Seeded orbit views for batched rendering: view v looks at the origin of
    scene v % n_scenes from azimuth ~ U(0, 2 pi), elevation ~ U(-0.4, 0.4) rad
    and distance ~ U(2.5, 5), with a resolution drawn uniformly from
    ``resolutions`` (fov 60 deg): view v draws 4 uniforms from stream
    (0xC1DE5EED00000000 + seed, D_VIEW, v): resolution floor(len u0), azimuth
    2 pi u1, elevation -0.4 + 0.8 u2, distance 2.5 + 2.5 u3.  Returns
    (view_scene, cameras, compositors)."""
    U = R.Streams(SEED_BASE + seed, R.D_VIEW, n_views).uniform(4)
    view_scene, cams, comps = [], [], []
    for v in range(n_views):
        W, H = resolutions[min(int(U[v, 0] * len(resolutions)), len(resolutions) - 1)]
        az = 2.0 * math.pi * U[v, 1]
        el = -0.4 + 0.8 * U[v, 2]
        r = 2.5 + 2.5 * U[v, 3]
        eye = (r * math.cos(el) * math.sin(az), r * math.sin(el), -r * math.cos(el) * math.cos(az))
        view_scene.append(v % n_scenes)
        cams.append(camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, W, H))
        comps.append(compositor(W, H, eye))
    return view_scene, cams, comps
