"""ctypes binding of the C ABI in include/cindergs.h (libcindergs_hip.so).

The library is built in-tree by ``cindergs_amd.build.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcindergs_hip.so")

# status codes (cgs_status)
CGS_OK = 0
# what a near/far network learns (cgs_mlp_set_far_kind)
CGS_FAR_SHARE = 0
CGS_FAR_SUMS = 1
CGS_ERR_RUNTIME = 1
CGS_ERR_NO_DEVICE = 2
CGS_ERR_UNSUPPORTED = 3
CGS_ERR_EXECUTION = 4
CGS_ERR_SHORT_READBACK = 5
CGS_ERR_INVALID_ARG = 6
CGS_ERR_FORMAT = 7

STATUS_NAMES = {
    CGS_ERR_RUNTIME: "Runtime",
    CGS_ERR_NO_DEVICE: "NoAdapter",
    CGS_ERR_UNSUPPORTED: "UnsupportedTarget",
    CGS_ERR_EXECUTION: "Execution",
    CGS_ERR_SHORT_READBACK: "ShortReadback",
    CGS_ERR_INVALID_ARG: "InvalidArgument",
    CGS_ERR_FORMAT: "Format",
}

# PlyError kinds in PlyLayout.error (include/cindergs.h)
PLY_ERRORS = {
    1: "MissingEndHeader", 2: "InvalidMagic", 3: "InvalidHeaderUtf8", 4: "UnsupportedFormat",
    5: "MissingFormat", 6: "MissingVertexElement", 7: "UnsupportedLayout",
    8: "UnsupportedListProperty", 9: "UnsupportedElement", 10: "MissingProperty",
    11: "InvalidPropertyType", 12: "UnsupportedShLayout", 13: "TruncatedPayload",
}


class PlyLayout(C.Structure):
    """cgs_ply_layout: PlyHeader (ply_loader.rs:125-140) as a decode table."""

    _fields_ = [
        ("vertex_count", C.c_uint64),
        ("payload_offset", C.c_uint64),
        ("vertex_stride", C.c_uint32),
        ("format", C.c_uint32),
        ("sh_degree", C.c_uint32),
        ("canonical", C.c_uint32),
        ("num_properties", C.c_uint32),
        ("error", C.c_uint32),
        ("error_detail", C.c_uint64 * 2),
        ("slot_offset", C.c_uint32 * 62),
        ("slot_type", C.c_uint8 * 62),
        ("pad", C.c_uint8 * 2),
    ]


class CameraUniforms(C.Structure):
    """CameraUniforms, 96 bytes (src/render/camera.rs:9-36)."""

    _fields_ = [
        ("view_matrix", C.c_float * 16),
        ("focal_x", C.c_float),
        ("focal_y", C.c_float),
        ("principal_x", C.c_float),
        ("principal_y", C.c_float),
        ("viewport_width", C.c_float),
        ("viewport_height", C.c_float),
        ("near_plane", C.c_float),
        ("pad", C.c_float),
    ]


class CompositorUniforms(C.Structure):
    """CompositorUniforms, 48 bytes (src/render/rasterizer/mod.rs:16-37)."""

    _fields_ = [
        ("screen_width", C.c_uint32),
        ("screen_height", C.c_uint32),
        ("tiles_x", C.c_uint32),
        ("tiles_y", C.c_uint32),
        ("background_color", C.c_float * 4),
        ("camera_position", C.c_float * 3),
        ("sh_degree", C.c_uint32),
    ]


class FrameParams(C.Structure):
    _fields_ = [
        ("strain_split_threshold", C.c_float),
        ("mip_filter_3d_var", C.c_float),
        ("min_opacity", C.c_float),
        ("min_radius_px", C.c_float),
        ("profile_stages", C.c_uint32),
        ("force_stage3", C.c_uint32),
    ]


class SortFrame(C.Structure):
    """cgs_sort_frame: the frame fields SortCache::take_if_reusable_for_frame keys on
    (src/render/pipeline.rs:119-145)."""
    _fields_ = [
        ("splat_count", C.c_uint64),
        ("focal_x", C.c_float),
        ("focal_y", C.c_float),
        ("split_threshold", C.c_float),
        ("mip_var", C.c_float),
    ]


class StageTimings(C.Structure):
    """GpuStageTimings (src/render/gpu_state.rs:513-534)."""

    _fields_ = [
        ("stage1_deform_ms", C.c_double),
        ("stage2_project_ms", C.c_double),
        ("stage3_bin_sort_ms", C.c_double),
        ("stage4_composite_ms", C.c_double),
        ("stage1_skipped", C.c_uint32),
        ("stage3_cached", C.c_uint32),
        ("used_hardware_timestamps", C.c_uint32),
        ("stage3_reruns", C.c_uint32),
        ("active_gaussians", C.c_uint64),
        ("surviving_gaussians", C.c_uint64),
        ("total_tile_entries", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    """cgs_tuning: launch shapes that leave every output bit-identical, and
    the tile size (16 x 16 the reference's; 8 x 8, 16 x 32, 32 x 16 render the
    same algorithm on that tile grid)."""

    _fields_ = [
        ("radix_partition_keys", C.c_uint32),
        ("radix_digit_bits", C.c_uint32),
        ("composite_batch", C.c_uint32),
        ("tile_order", C.c_uint32),
        ("tile_sort", C.c_uint32),
        ("tile_w", C.c_uint32),
        ("tile_h", C.c_uint32),
    ]

    def __init__(self, radix_partition_keys=16384, radix_digit_bits=8, composite_batch=64,
                 tile_order=1, tile_sort=0, tile_w=16, tile_h=16):
        super().__init__(radix_partition_keys, radix_digit_bits, composite_batch, tile_order,
                         tile_sort, tile_w, tile_h)

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(C.Structure):
    """cgs_batch_stats: one cgs_execute_views_batched call."""

    _fields_ = [
        ("views", C.c_uint32),
        ("lanes", C.c_uint32),
        ("graph_built", C.c_uint32),
        ("reruns", C.c_uint32),
        ("total_tile_entries", C.c_uint64),
        ("gpu_ms", C.c_double),
        ("presort_views", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


assert C.sizeof(CameraUniforms) == 96
assert C.sizeof(CompositorUniforms) == 48

_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_int = C.c_int
_f = C.c_float

# (name, argtypes) of every entry point declared in include/cindergs.h
SIGNATURES = {
    "cgs_create": [_int, C.POINTER(_vp)],
    "cgs_destroy": [_vp],
    "cgs_last_error": [],
    "cgs_device_name": [_vp, C.c_char_p, C.c_size_t],
    "cgs_synchronize": [_vp],
    "cgs_invalidate": [_vp],
    "cgs_invalidate_stage3": [_vp],
    "cgs_sync_scene": [_vp, _vp, _vp, _vp, _vp, _u64],
    "cgs_sync_scene_f16": [_vp, _vp, _vp, _vp, _vp, _u32, _u64],
    "cgs_sync_cages": [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _u64],
    "cgs_execute_frame": [_vp, C.POINTER(CameraUniforms), C.POINTER(CompositorUniforms),
                          C.POINTER(FrameParams), C.POINTER(StageTimings)],
    "cgs_readback_rgba8": [_vp, _vp, _u32, _u32],
    "cgs_framebuffer_device_ptr": [_vp, C.POINTER(_vp), C.POINTER(_u32)],
    "cgs_kernel_times": [_vp, C.c_char_p, C.POINTER(C.c_double), _u32, C.POINTER(_u32)],
    "cgs_set_tuning": [_vp, C.POINTER(Tuning)],
    "cgs_get_tuning": [_vp, C.POINTER(Tuning)],
    "cgs_autotune": [_vp, C.POINTER(CameraUniforms), C.POINTER(CompositorUniforms),
                     C.POINTER(FrameParams), _u32, _u32, C.POINTER(Tuning), C.POINTER(C.c_double),
                     _u32],
    "cgs_tuning_candidates": [],
    "cgs_tuning_candidate": [_u32, C.POINTER(Tuning)],
    "cgs_batch_set_tuning": [_vp, C.POINTER(Tuning)],
    "cgs_batch_set_shared_projection": [_vp, _u32],
    "cgs_batch_create": [_int, _u32, C.POINTER(_vp)],
    "cgs_batch_destroy": [_vp],
    "cgs_execute_views_batched": [_vp, _vp, _u32, _vp, _vp, _vp, _u32, C.POINTER(FrameParams),
                                  C.POINTER(BatchStats)],
    "cgs_batch_framebuffer": [_vp, _u32, C.POINTER(_vp), C.POINTER(_u32)],
    "cgs_batch_readback_rgba8": [_vp, _u32, _vp, _u32, _u32],
    "cgs_batch_view_counts": [_vp, _u32, C.POINTER(_u32), C.POINTER(_u32)],
    "cgs_read_projected": [_vp, _vp, _vp, _u64],
    "cgs_read_stage3": [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_u64)],
    "cgs_set_capture": [_vp, _u32],
    "cgs_read_deformed": [_vp, _vp, _vp, _vp, _vp, C.POINTER(_u32)],
    "cgs_read_tet_transforms": [_vp, _vp, _vp, _vp],
    "cgs_export_deformed_scene": [_vp, _u32, _vp, _vp, _vp, _vp, _u64, C.POINTER(_u64)],
    "cgs_sort_cache_create": [C.POINTER(_vp)],
    "cgs_sort_cache_destroy": [_vp],
    "cgs_sort_cache_can_reuse": [_vp, _vp, _vp, _u64, _u32, _u32, C.POINTER(_u32)],
    "cgs_sort_cache_take_if_reusable": [_vp, _vp, _vp, _u64, _u32, _u32, C.POINTER(_u32)],
    "cgs_sort_cache_take_if_reusable_for_frame": [_vp, _vp, _vp, _u64, _u32, _u32, _vp,
                                                  C.POINTER(_u32)],
    "cgs_sort_cache_commit": [_vp, _vp, _vp, _u64, _u32, _u32],
    "cgs_sort_cache_commit_for_frame": [_vp, _vp, _vp, _u64, _u32, _u32, _vp],
    "cgs_sort_cache_invalidate": [_vp],
    "cgs_sort_cache_reuse_count": [_vp, C.POINTER(_u64)],
    "cgs_set_sort_hysteresis": [_vp, _u32],
    "cgs_sort_hysteresis_reuse_count": [_vp, C.POINTER(_u64)],
    "cgs_bind_tets": [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _u64, _vp, _vp, _vp],
    "cgs_mvc_bind_exact": [_vp, _vp, _u32, _vp, _u32, _vp, _u64, _vp, _vp, _vp],
    "cgs_mvc_far_share": [_vp, _vp, _u32, _vp, _u32, _f, _f, _vp, _u64, _vp],
    "cgs_mvc_nearfar_bind": [_vp, _vp, _u32, _vp, _u32, _f, _f, _vp, _u64, _vp, _vp, _vp, _vp, _vp,
                             _vp],
    "cgs_mlp_set_near_field": [_vp, _f, _f],
    "cgs_mlp_set_far_kind": [_vp, _u32],
    "cgs_mvc_far_targets": [_vp, _vp, _u32, _vp, _u32, _f, _f, _vp, _u64, _u32, _vp],
    "cgs_mvc_nearfar_bind_kind": [_vp, _vp, _u32, _vp, _u32, _f, _f, _vp, _u64, _vp, _vp, _u32,
                                  _vp, _vp, _vp, _vp],
    "cgs_mvc_winding": [_vp, _vp, _u32, _vp, _u32, _vp, _u64, _vp],
    "cgs_mvc_deform": [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u64, _vp, _vp],
    "cgs_mvc_deform_scene": [_vp, _vp, _vp, _vp, _vp, _u32, _u64, _vp, _vp, _vp, _vp, _u32,
                             C.c_float, C.c_float, _vp, _vp, _vp, _vp, C.POINTER(_u64)],
    "cgs_ply_parse_header": [_vp, _u64, _vp],
    "cgs_ply_load": [_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp],
    "cgs_activate_raw": [_vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp],
    "cgs_quantize_scene": [_vp, _vp, _vp, _u64, _u32, _vp, _vp, _vp],
    "cgs_export_raw": [_vp, _vp, _vp, _vp, _vp, _u64, _vp],
    "cgs_mlp_create": [_vp, _u32, C.c_float, _u64, C.POINTER(_vp)],
    "cgs_mlp_destroy": [_vp],
    "cgs_mlp_param_count": [_vp],
    "cgs_mlp_layout": [_vp, _vp, _vp, _vp],
    "cgs_mlp_get_params": [_vp, _vp, _vp],
    "cgs_mlp_set_params": [_vp, _vp, _vp],
    "cgs_mlp_train": [_vp, _vp, _vp, _u64, _u32, _u32, C.c_float, _vp],
    "cgs_mlp_gradient": [_vp, _vp, _vp, _u32, _vp, _vp],
    "cgs_mlp_infer": [_vp, _vp, _u64, _vp, _vp],
    "cgs_mlp_bind": [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _u64, C.c_float, _vp, _vp, _vp,
                     C.POINTER(_u64)],
    "cgs_load_projected": [_vp, _vp, _vp, _u64],
    "cgs_run_stage3": [_vp, C.POINTER(CameraUniforms), C.POINTER(CompositorUniforms),
                       C.POINTER(_u64)],
    "cgs_run_stage4": [_vp, C.POINTER(CompositorUniforms)],
    "cgs_malloc": [_vp, _u64, C.POINTER(_vp)],
    "cgs_free": [_vp, _vp],
    "cgs_memcpy_h2d": [_vp, _vp, _vp, _u64],
    "cgs_memcpy_d2h": [_vp, _vp, _vp, _u64],
    "cgs_memcpy_d2d": [_vp, _vp, _vp, _u64],
    "cgs_exclusive_scan_u32": [_vp, _vp, _vp, _u64, C.POINTER(_u32)],
    "cgs_exclusive_scan_u32_noalloc": [_vp, _vp, _vp, _u64, C.POINTER(_u32)],
    "cgs_radix_sort_u32_pairs": [_vp, _vp, _vp, _u64, _u32],
    "cgs_radix_sort_u64_pairs": [_vp, _vp, _vp, _vp, _u64, _u32],
}

_lib = None


def load() -> C.CDLL:
    """Load libcindergs_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        if name == "cgs_last_error":
            fn.restype = C.c_char_p
        elif name in ("cgs_destroy", "cgs_mlp_destroy", "cgs_batch_destroy",
                      "cgs_sort_cache_destroy"):
            fn.restype = None
        elif name == "cgs_mlp_param_count":
            fn.restype = _u64
        elif name == "cgs_tuning_candidates":
            fn.restype = _u32
        else:
            fn.restype = C.c_int
    _lib = lib
    return lib


class GpuError(RuntimeError):
    """GpuError / GpuBridgeError of the reference (src/render/gpu.rs:14-30)."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = STATUS_NAMES.get(code, f"Status{code}")
        super().__init__(f"{self.kind}: {message}")


def check(code: int) -> None:
    if code != CGS_OK:
        msg = load().cgs_last_error()
        raise GpuError(code, msg.decode() if msg else "")
