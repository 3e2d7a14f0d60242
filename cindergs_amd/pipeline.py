"""Host-side mirror of the reference's GPU pipeline API over the C ABI.

Names follow src/render/gpu_state.rs: ``GpuPipelineState`` owns the device
context (scene residency = ``GpuSceneState``, buffers = ``GpuFrameArena``) and
``execute_frame_gpu`` runs Stages 1-4.  Stage primitives follow
src/render/kernels/{scan,radix_sort}.rs.  Arrays are numpy arrays in the
reference layouts (GaussianSceneSoa, CameraUniforms, attrs[10]).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (CameraUniforms, CompositorUniforms, FrameParams, GpuError, StageTimings,
                   check)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def padded_rgba8_row_u32(width: int) -> int:
    """Row stride in u32 pixels, 256-byte aligned (gpu_state.rs:25-27)."""
    return (width + 63) // 64 * 64


@dataclass
class GaussianSceneSoa:
    """src/scene/soa.rs:30-45 (f32, activated opacities, 48 SH coefficients)."""

    positions: np.ndarray      # [n,3] f32
    covariances_3d: np.ndarray  # [n,6] f32 (xx,xy,xz,yy,yz,zz)
    opacities: np.ndarray      # [n] f32
    sh_coeffs: np.ndarray      # [n,48] f32

    @property
    def count(self) -> int:
        return int(self.opacities.shape[0])

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.covariances_3d = np.ascontiguousarray(self.covariances_3d, dtype=np.float32).reshape(-1, 6)
        self.opacities = np.ascontiguousarray(self.opacities, dtype=np.float32).reshape(-1)
        self.sh_coeffs = np.ascontiguousarray(self.sh_coeffs, dtype=np.float32).reshape(-1, 48)


@dataclass
class QuantizedGaussianSceneSoa:
    """src/scene/soa.rs:429-500: f16-pair covariances and degree-trimmed f16 SH."""

    positions: np.ndarray              # [n,3] f32
    covariances_f16_packed: np.ndarray  # [n,3] u32
    opacities: np.ndarray              # [n] f32
    sh_coeffs_f16_packed: np.ndarray    # [n, words] u32
    sh_degree: int
    octahedral_rotations: Optional[np.ndarray] = None  # [n] u32 (soa.rs:369-402)

    def total_bytes(self) -> int:
        """Quantised footprint (soa.rs:581-589)."""
        rot = 0 if self.octahedral_rotations is None else self.octahedral_rotations.size * 4
        return (self.positions.size + self.covariances_f16_packed.size + self.opacities.size
                + self.sh_coeffs_f16_packed.size) * 4 + rot

    def unquantized_bytes(self) -> int:
        return self.count * (3 + 6 + 1 + 48) * 4

    def compression_ratio(self) -> float:
        q = self.total_bytes()
        return 1.0 if q == 0 else self.unquantized_bytes() / q

    @property
    def count(self) -> int:
        return int(self.opacities.shape[0])


@dataclass
class TetCage:
    """A flattened CageScene (src/deformation/mod.rs:420-464): global tets,
    their inverse rest edge matrices, current vertices and per-splat bindings."""

    elements: np.ndarray   # [T,4] u32
    inv_dm: np.ndarray     # [T,9] f32 column-major
    vertices: np.ndarray   # [V,3] f32
    bind_tet: np.ndarray   # [n] u32, 0xFFFFFFFF = unbound
    bind_weights: np.ndarray  # [n,4] f32


class GpuPipelineState:
    """Persistent GPU-resident 4-stage pipeline (gpu_state.rs:536-1347)."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.cgs_create(device, C.byref(h)))
        self._h = h
        self._lib = lib
        self._keep = []  # host arrays whose pointers feed the fingerprint
        self.width = 0
        self.height = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- residency --------------------------------------------------------
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        check(self._lib.cgs_device_name(self._h, buf, 256))
        return buf.value.decode()

    def invalidate(self):
        check(self._lib.cgs_invalidate(self._h))

    def invalidate_stage3(self):
        check(self._lib.cgs_invalidate_stage3(self._h))

    def sync_scene(self, scene):
        if isinstance(scene, QuantizedGaussianSceneSoa):
            check(self._lib.cgs_sync_scene_f16(
                self._h, _ptr(scene.positions), _ptr(scene.covariances_f16_packed),
                _ptr(scene.opacities), _ptr(scene.sh_coeffs_f16_packed), scene.sh_degree,
                scene.count))
        else:
            check(self._lib.cgs_sync_scene(
                self._h, _ptr(scene.positions), _ptr(scene.covariances_3d), _ptr(scene.opacities),
                _ptr(scene.sh_coeffs), scene.count))
        self._keep = [scene]

    def sync_cages(self, cage: Optional[TetCage], count: int):
        if cage is None:
            check(self._lib.cgs_sync_cages(self._h, None, None, 0, None, 0, None, None, count))
            return
        check(self._lib.cgs_sync_cages(
            self._h, _ptr(cage.elements), _ptr(cage.inv_dm), cage.elements.shape[0],
            _ptr(cage.vertices), cage.vertices.shape[0], _ptr(cage.bind_tet),
            _ptr(cage.bind_weights), count))

    def sync_scene_and_cages(self, scene, cages: Optional[TetCage] = None):
        """GpuSceneState::sync_scene_and_cages (gpu_state.rs:182-343)."""
        self.sync_scene(scene)
        self.sync_cages(cages, scene.count)

    # ---- frame -------------------------------------------------------------
    def execute_frame_gpu(self, scene, cages, camera: CameraUniforms,
                          compositor: CompositorUniforms, strain_split_threshold: float = 0.0,
                          mip_filter_3d_var: float = 0.0, min_opacity: float = 1.0 / 255.0,
                          min_radius_px: float = 0.5, profile_stages: bool = False,
                          force_stage3: bool = False) -> StageTimings:
        """gpu_state.rs:613-626 (scene/cages synchronised first, as the reference)."""
        if scene is not None:
            self.sync_scene_and_cages(scene, cages)
        p = FrameParams(strain_split_threshold, mip_filter_3d_var, min_opacity, min_radius_px,
                        int(profile_stages), int(force_stage3))
        t = StageTimings()
        check(self._lib.cgs_execute_frame(self._h, C.byref(camera), C.byref(compositor),
                                          C.byref(p), C.byref(t)))
        self.width = max(1, compositor.screen_width)
        self.height = max(1, compositor.screen_height)
        return t

    def execute_frame_resident(self, camera, compositor, params: FrameParams,
                               timings: StageTimings) -> None:
        """Frame on the already-synchronised scene (no host work besides the call)."""
        check(self._lib.cgs_execute_frame(self._h, C.byref(camera), C.byref(compositor),
                                          C.byref(params), C.byref(timings)))

    # ---- launch tuning ------------------------------------------------------
    def set_tuning(self, tuning: "_lib.Tuning"):
        check(self._lib.cgs_set_tuning(self._h, C.byref(tuning)))

    def get_tuning(self) -> "_lib.Tuning":
        t = _lib.Tuning()
        check(self._lib.cgs_get_tuning(self._h, C.byref(t)))
        return t

    def autotune(self, camera, compositor, params: FrameParams, reps: int = 3,
                 tile_shapes: bool = False):
        """Time the tuning candidates of the current tile size (of every size
        with tile_shapes) on this view, each checked to give the same frame as
        the others of its size, and keep the fastest.  Returns (best,
        [(candidate, ms)]) with ms = -1 for the candidates not swept."""
        n = tuning_candidates_count()
        ms = (C.c_double * n)()
        best = _lib.Tuning()
        flags = 1 if tile_shapes else 0  # CGS_TUNE_TILE_SHAPES
        check(self._lib.cgs_autotune(self._h, C.byref(camera), C.byref(compositor),
                                     C.byref(params), reps, flags, C.byref(best), ms, n))
        return best, [(tuning_candidate(i), ms[i]) for i in range(n)]

    def kernel_times(self):
        """[(kernel name, ms)] of the last frame run with profile_stages=2."""
        names = C.create_string_buffer(32 * 64)
        ms = (C.c_double * 64)()
        n = C.c_uint32()
        check(self._lib.cgs_kernel_times(self._h, names, ms, 64, C.byref(n)))
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), ms[i]) for i in range(n.value)]

    def readback_framebuffer_rgba8(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.uint8)
        check(self._lib.cgs_readback_rgba8(self._h, _ptr(out), width, height))
        return out

    def framebuffer_device_ptr(self):
        p = C.c_void_p()
        row = C.c_uint32()
        check(self._lib.cgs_framebuffer_device_ptr(self._h, C.byref(p), C.byref(row)))
        return p.value, row.value

    # ---- stage access ------------------------------------------------------
    def set_capture(self, enable: bool):
        check(self._lib.cgs_set_capture(self._h, int(enable)))

    def read_projected(self, n: int):
        attrs = np.zeros((n, 10), dtype=np.float32)
        ext = np.zeros(n, dtype=np.uint32)
        check(self._lib.cgs_read_projected(self._h, _ptr(attrs), _ptr(ext), n))
        return attrs, ext

    def read_stage3(self, max_gaussians: int, total_tiles: int, with_counts: bool = False):
        n = C.c_uint64()
        check(self._lib.cgs_read_stage3(self._h, None, None, None, None, None, C.byref(n)))
        E = n.value
        keys = np.zeros(max(E, 1), dtype=np.uint32)
        ids = np.zeros(max(E, 1), dtype=np.uint32)
        ranges = np.zeros(total_tiles * 2, dtype=np.uint32)
        counts = np.zeros(max_gaussians, dtype=np.uint32) if with_counts else None
        offsets = np.zeros(max_gaussians, dtype=np.uint32) if with_counts else None
        check(self._lib.cgs_read_stage3(self._h, _ptr(counts), _ptr(offsets), _ptr(keys),
                                        _ptr(ids), _ptr(ranges), C.byref(n)))
        out = {"keys": keys[:E], "ids": ids[:E], "ranges": ranges.reshape(-1, 2), "n": E}
        if with_counts:
            out["counts"] = counts
            out["offsets"] = offsets
        return out

    def read_deformed(self, n_rest: int):
        cnt = C.c_uint32()
        check(self._lib.cgs_read_deformed(self._h, None, None, None, None, C.byref(cnt)))
        m = n_rest + cnt.value
        pos = np.zeros((m, 3), dtype=np.float32)
        cov = np.zeros((m, 6), dtype=np.float32)
        opa = np.zeros(m, dtype=np.float32)
        meta = np.zeros((max(cnt.value, 1), 2), dtype=np.uint32)
        check(self._lib.cgs_read_deformed(self._h, _ptr(pos), _ptr(cov), _ptr(opa), _ptr(meta),
                                          C.byref(cnt)))
        return pos, cov, opa, meta[: cnt.value]

    def export_deformed_scene(self, sh_degree: int = 3) -> GaussianSceneSoa:
        """The deformed scene of the last frame with cages as deform_scene_cpu returns
        it (cpu_reference.rs:97-137, 166-182, 394-402), built on the device: Stage 1's
        positions, covariances and opacities, and each record's source SH rotated
        by its tet's ShRotator (sh_rotation.rs:396-446)."""
        cnt = C.c_uint64()
        check(self._lib.cgs_export_deformed_scene(self._h, sh_degree, None, None, None, None,
                                                  0, C.byref(cnt)))
        m = cnt.value
        shapes = ((m, 3), (m, 6), (m,), (m, 48))
        if m == 0:
            return GaussianSceneSoa(*[np.zeros(s, np.float32) for s in shapes])
        bufs = [self.to_device(np.zeros(s, np.float32)) for s in shapes]
        try:
            check(self._lib.cgs_export_deformed_scene(self._h, sh_degree,
                                                      *[b.ptr for b in bufs], m,
                                                      C.byref(cnt)))
            out = [b.numpy() for b in bufs]
        finally:
            for b in bufs:
                b.free()
        return GaussianSceneSoa(*out)

    def set_sort_hysteresis(self, enable: bool):
        """The CPU fallback frame's SortCache hysteresis (pipeline.rs:539-600) inside
        execute_frame_gpu: reuse the previous tile order while the camera and the
        cages move by at most 1e-5."""
        check(self._lib.cgs_set_sort_hysteresis(self._h, 1 if enable else 0))

    @property
    def sort_reuse_count(self) -> int:
        n = C.c_uint64()
        check(self._lib.cgs_sort_hysteresis_reuse_count(self._h, C.byref(n)))
        return n.value

    def read_tet_transforms(self, num_tets: int):
        J = np.zeros((num_tets, 9), dtype=np.float32)
        R = np.zeros((num_tets, 9), dtype=np.float32)
        s = np.zeros(num_tets, dtype=np.float32)
        check(self._lib.cgs_read_tet_transforms(self._h, _ptr(J), _ptr(R), _ptr(s)))
        return J, R, s

    def bind_tets(self, rest_vertices: np.ndarray, elements: np.ndarray, inv_dm: np.ndarray,
                  positions: np.ndarray, bounds=None):
        """TetMesh::bind_gaussians(_with_bounds) (mod.rs:263-323) on the GPU.
        Returns (tet[n] u32, 0xFFFFFFFF = unbound; weights[n,4] f32)."""
        v = np.ascontiguousarray(rest_vertices, dtype=np.float32)
        e = np.ascontiguousarray(elements, dtype=np.uint32)
        m = np.ascontiguousarray(inv_dm, dtype=np.float32)
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        b = None
        if bounds is not None:
            b = np.ascontiguousarray(np.concatenate([np.asarray(bounds[0], np.float32),
                                                     np.asarray(bounds[1], np.float32)]))
        tet = np.zeros(p.shape[0], np.uint32)
        w = np.zeros((p.shape[0], 4), np.float32)
        check(self._lib.cgs_bind_tets(self._h, _ptr(v), v.shape[0], _ptr(e), _ptr(m),
                                      e.shape[0], _ptr(p), p.shape[0], _ptr(b), _ptr(tet),
                                      _ptr(w)))
        return tet, w

    def load_projected(self, attrs10: np.ndarray, extents: np.ndarray):
        a = np.ascontiguousarray(attrs10, dtype=np.float32)
        e = np.ascontiguousarray(extents, dtype=np.uint32)
        check(self._lib.cgs_load_projected(self._h, _ptr(a), _ptr(e), e.shape[0]))

    def run_stage3(self, camera: CameraUniforms, compositor: CompositorUniforms) -> int:
        n = C.c_uint64()
        check(self._lib.cgs_run_stage3(self._h, C.byref(camera), C.byref(compositor), C.byref(n)))
        self.width = max(1, compositor.screen_width)
        self.height = max(1, compositor.screen_height)
        return n.value

    def run_stage4(self, compositor: CompositorUniforms):
        check(self._lib.cgs_run_stage4(self._h, C.byref(compositor)))

    def synchronize(self):
        check(self._lib.cgs_synchronize(self._h))

    # ---- device buffers and primitives ------------------------------------
    def to_device(self, a: np.ndarray) -> "DeviceBuffer":
        return DeviceBuffer(self, a)

    def exclusive_scan_total(self, d_in: "DeviceBuffer", d_out: "DeviceBuffer", n: int) -> int:
        """gpu_exclusive_scan_total (scan.rs:312-333)."""
        t = C.c_uint32()
        check(self._lib.cgs_exclusive_scan_u32(self._h, d_in.ptr, d_out.ptr, n, C.byref(t)))
        return t.value

    def exclusive_scan_total_noalloc(self, d_in: "DeviceBuffer", d_out: "DeviceBuffer", n: int,
                                     want_total: bool = True):
        """The scan without any allocation (cgs_exclusive_scan_u32_noalloc):
        the total, or None (and an asynchronous call) with want_total False."""
        t = C.c_uint32()
        check(self._lib.cgs_exclusive_scan_u32_noalloc(self._h, d_in.ptr if d_in else None,
                                                       d_out.ptr if d_out else None, n,
                                                       C.byref(t) if want_total else None))
        return t.value if want_total else None

    def radix_sort_u32_pairs(self, keys: "DeviceBuffer", vals: "DeviceBuffer", n: int,
                             sort_bits: int):
        """gpu_radix_sort_u32_pairs (radix_sort.rs:292-368), in place."""
        check(self._lib.cgs_radix_sort_u32_pairs(self._h, keys.ptr, vals.ptr, n, sort_bits))

    def radix_sort_u64_pairs(self, tiles: "DeviceBuffer", depths: "DeviceBuffer",
                             ids: "DeviceBuffer", n: int, tile_bits: int):
        """gpu_radix_sort_u64_pairs (radix_sort.rs:377-477), in place."""
        check(self._lib.cgs_radix_sort_u64_pairs(self._h, tiles.ptr, depths.ptr, ids.ptr, n,
                                                 tile_bits))


def _view16(view) -> np.ndarray:
    v = np.ascontiguousarray(view, dtype=np.float32).reshape(-1)
    if v.shape[0] != 16:
        raise ValueError("a view matrix has 16 floats")
    return v


def _cage_xyz(cage_state):
    c = np.ascontiguousarray(cage_state, dtype=np.float32).reshape(-1, 3)
    return c, c.shape[0]


class SortCache:
    """SortCache (src/render/pipeline.rs:66-193) over the C ABI's cgs_sort_cache.
    The binning (any object: the reference's TileBinningOutputs) is held here;
    the native object decides reuse and counts reuses."""

    def __init__(self):
        self._lib = _lib.load()
        p = C.c_void_p()
        check(self._lib.cgs_sort_cache_create(C.byref(p)))
        self._h = p
        self.binning = None

    def close(self):
        if self._h:
            self._lib.cgs_sort_cache_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _frame(splat_count, focal_x, focal_y, split_threshold, mip_var):
        return _lib.SortFrame(int(splat_count), float(focal_x), float(focal_y),
                              float(split_threshold), float(mip_var))

    def can_reuse(self, view_matrix, cage_state, width: int, height: int) -> bool:
        v, (c, n) = _view16(view_matrix), _cage_xyz(cage_state)
        r = C.c_uint32()
        check(self._lib.cgs_sort_cache_can_reuse(self._h, _ptr(v), _ptr(c), n, width, height,
                                                 C.byref(r)))
        return bool(r.value)

    def _take(self, reused: int):
        b, self.binning = self.binning, None
        return b if reused else None

    def take_if_reusable(self, view_matrix, cage_state, width: int, height: int):
        v, (c, n) = _view16(view_matrix), _cage_xyz(cage_state)
        r = C.c_uint32()
        check(self._lib.cgs_sort_cache_take_if_reusable(self._h, _ptr(v), _ptr(c), n, width,
                                                        height, C.byref(r)))
        return self._take(r.value)

    def take_if_reusable_for_frame(self, view_matrix, cage_state, width: int, height: int,
                                   splat_count: int, focal_x: float, focal_y: float,
                                   split_threshold: float, mip_var: float):
        v, (c, n) = _view16(view_matrix), _cage_xyz(cage_state)
        f = self._frame(splat_count, focal_x, focal_y, split_threshold, mip_var)
        r = C.c_uint32()
        check(self._lib.cgs_sort_cache_take_if_reusable_for_frame(
            self._h, _ptr(v), _ptr(c), n, width, height, C.byref(f), C.byref(r)))
        return self._take(r.value)

    def commit(self, view_matrix, cage_state, width: int, height: int, binning):
        v, (c, n) = _view16(view_matrix), _cage_xyz(cage_state)
        check(self._lib.cgs_sort_cache_commit(self._h, _ptr(v), _ptr(c), n, width, height))
        self.binning = binning

    def commit_for_frame(self, view_matrix, cage_state, width: int, height: int,
                         splat_count: int, focal_x: float, focal_y: float,
                         split_threshold: float, mip_var: float, binning):
        v, (c, n) = _view16(view_matrix), _cage_xyz(cage_state)
        f = self._frame(splat_count, focal_x, focal_y, split_threshold, mip_var)
        check(self._lib.cgs_sort_cache_commit_for_frame(self._h, _ptr(v), _ptr(c), n, width,
                                                        height, C.byref(f)))
        self.binning = binning

    def invalidate(self):
        check(self._lib.cgs_sort_cache_invalidate(self._h))
        self.binning = None

    @property
    def reuse_count(self) -> int:
        n = C.c_uint64()
        check(self._lib.cgs_sort_cache_reuse_count(self._h, C.byref(n)))
        return n.value


class DeviceBuffer:
    """A device allocation owned by a GpuPipelineState's device."""

    def __init__(self, state: GpuPipelineState, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.state = state
        self.dtype = a.dtype
        self.shape = a.shape
        self.nbytes = max(int(a.nbytes), 4)
        p = C.c_void_p()
        check(state._lib.cgs_malloc(state._h, self.nbytes, C.byref(p)))
        self.ptr = p
        if a.nbytes:
            check(state._lib.cgs_memcpy_h2d(state._h, p, _ptr(a), a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            check(self.state._lib.cgs_memcpy_d2h(self.state._h, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.state._h:
            check(self.state._lib.cgs_free(self.state._h, self.ptr))
        self.ptr = None


def tuning_candidates_count() -> int:
    return int(_lib.load().cgs_tuning_candidates())


def tuning_candidate(i: int) -> "_lib.Tuning":
    t = _lib.Tuning()
    check(_lib.load().cgs_tuning_candidate(i, C.byref(t)))
    return t


class ViewBatch:
    """Many views of several resident scenes from one captured HIP graph.

    The reference renders one view per ``execute_frame_gpu`` call
    (gpu_state.rs:613-1243) in a per-view loop of its callers
    (benchmark.rs:231-366, headless.rs:140-210); this renders a list of views
    in one call.  Each view's pixels equal ``execute_frame_gpu`` on the same
    scene, camera and parameters.  ``lanes`` views run concurrently.
    """

    def __init__(self, device: int = 0, lanes: int = 4):
        lib = _lib.load()
        h = C.c_void_p()
        check(lib.cgs_batch_create(device, lanes, C.byref(h)))
        self._h = h
        self._lib = lib
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cgs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, tuning: "_lib.Tuning"):
        check(self._lib.cgs_batch_set_tuning(self._h, C.byref(tuning)))

    def set_shared_projection(self, enable: bool):
        """Stage 2 of all views of a scene in one pass over the scene (default)."""
        check(self._lib.cgs_batch_set_shared_projection(self._h, int(enable)))

    def execute(self, states, view_scene, cameras, compositors, params: FrameParams):
        """Render view v of ``states[view_scene[v]]`` (synchronised GpuPipelineStates)."""
        n = len(cameras)
        if len(compositors) != n or len(view_scene) != n:
            raise ValueError("cameras, compositors and view_scene differ in length")
        scenes = (C.c_void_p * max(1, len(states)))(*[st._h for st in states])
        vs = np.ascontiguousarray(np.asarray(view_scene, dtype=np.uint32))
        cams = (CameraUniforms * max(1, n))(*cameras)
        comps = (CompositorUniforms * max(1, n))(*compositors)
        stats = _lib.BatchStats()
        check(self._lib.cgs_execute_views_batched(
            self._h, scenes, len(states), _ptr(vs), cams, comps, n, C.byref(params),
            C.byref(stats)))
        self.sizes = [(max(1, c.screen_width), max(1, c.screen_height)) for c in compositors]
        return stats

    def readback_rgba8(self, view: int) -> np.ndarray:
        w, h = self.sizes[view]
        out = np.empty((h, w, 4), dtype=np.uint8)
        check(self._lib.cgs_batch_readback_rgba8(self._h, view, _ptr(out), w, h))
        return out

    def framebuffer_device_ptr(self, view: int):
        p = C.c_void_p()
        row = C.c_uint32()
        check(self._lib.cgs_batch_framebuffer(self._h, view, C.byref(p), C.byref(row)))
        return p.value, row.value

    def view_counts(self, view: int):
        """(tile entries, surviving splats) of a view of the last batch."""
        e = C.c_uint32()
        s = C.c_uint32()
        check(self._lib.cgs_batch_view_counts(self._h, view, C.byref(e), C.byref(s)))
        return e.value, s.value


__all__ = [
    "CameraUniforms", "CompositorUniforms", "FrameParams", "StageTimings", "GpuError",
    "GaussianSceneSoa", "QuantizedGaussianSceneSoa", "TetCage", "GpuPipelineState",
    "DeviceBuffer", "ViewBatch", "padded_rgba8_row_u32", "tuning_candidate",
    "tuning_candidates_count",
]
