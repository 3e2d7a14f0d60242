"""Seeded synthetic scenes and cameras for the benchmark workloads.

The reference has no generator for these distributions; the specification
here is ours:
  * RNG: numpy PCG64 seeded with 0xC1DE5EED0000000k for workload k.
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
"""
from __future__ import annotations

import math

import numpy as np

from .pipeline import (CameraUniforms, CompositorUniforms, GaussianSceneSoa,
                       QuantizedGaussianSceneSoa)

SEED_BASE = 0xC1DE5EED00000000
CONFIG2_SPLATS = 1_073_930


def _rng(k: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED_BASE + k))


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


def _attributes(rng: np.random.Generator, n: int, means: np.ndarray) -> GaussianSceneSoa:
    log_s = rng.normal(-4.5, 0.5, size=(n, 3))
    s = np.maximum(np.exp(log_s), 1e-4)
    q = rng.normal(0.0, 1.0, size=(n, 4))
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    R = quat_to_mat(q)
    M = R * s[:, None, :]
    S = M @ np.transpose(M, (0, 2, 1))
    cov = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], 1)
    opa = 1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.5, size=n)))
    sh = np.empty((n, 48))
    sh[:, :3] = rng.normal(0.0, 0.5, size=(n, 3))
    sh[:, 3:] = rng.normal(0.0, 0.05, size=(n, 45))
    return GaussianSceneSoa(means.astype(np.float32), cov.astype(np.float32),
                            opa.astype(np.float32), sh.astype(np.float32))


def scene_gaussian_ball(n: int = 100_000, k: int = 1) -> GaussianSceneSoa:
    """100k-splat scene: means ~ N(0, 0.5^2 I3)."""
    rng = _rng(k)
    means = rng.normal(0.0, 0.5, size=(n, 3))
    return _attributes(rng, n, means)


def scene_clusters(n: int = CONFIG2_SPLATS, k: int = 2, clusters: int = 32,
                   background: float = 0.10) -> GaussianSceneSoa:
    """Cluster-mixture scene (1,073,930 splats by default)."""
    rng = _rng(k)
    n_bg = int(round(n * background))
    n_cl = n - n_bg
    centres = rng.uniform(-3.0, 3.0, size=(clusters, 3))
    radii = rng.uniform(0.05, 0.6, size=(clusters, 3))
    qc = rng.normal(0.0, 1.0, size=(clusters, 4))
    qc /= np.linalg.norm(qc, axis=1, keepdims=True)
    Rc = quat_to_mat(qc)
    which = rng.integers(0, clusters, size=n_cl)
    local = rng.normal(0.0, 1.0, size=(n_cl, 3)) * radii[which]
    pts = np.einsum("nij,nj->ni", Rc[which], local) + centres[which]
    bg = rng.uniform(-4.0, 4.0, size=(n_bg, 3))
    means = np.concatenate([pts, bg], 0)
    return _attributes(rng, n, means)


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
    ``resolutions`` (fov 60 deg).  Returns (view_scene, cameras, compositors)."""
    rng = _rng(seed)
    view_scene, cams, comps = [], [], []
    for v in range(n_views):
        W, H = resolutions[int(rng.integers(0, len(resolutions)))]
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = rng.uniform(-0.4, 0.4)
        r = rng.uniform(2.5, 5.0)
        eye = (r * math.cos(el) * math.sin(az), r * math.sin(el), -r * math.cos(el) * math.cos(az))
        view_scene.append(v % n_scenes)
        cams.append(camera_look_at(eye, (0, 0, 0), (0, 1, 0), 60.0, W, H))
        comps.append(compositor(W, H, eye))
    return view_scene, cams, comps
