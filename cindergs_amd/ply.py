"""Scene ingest and quantised storage on the GPU.

Host mirror of PlyLoader (src/scene/ply_loader.rs:601-716) and of the
quantisation / export of GaussianSceneSoa (src/scene/soa.rs:294-328,
454-500) over the C ABI:

* ``parse_header`` runs the library's header parser (cgs_ply_parse_header)
  and raises ``PlyError`` with the reference's error kind and message;
* ``PlyLoader.load_from_bytes`` / ``load_file`` copy the vertex payload to the
  device as is and decode + activate every record there (cgs_ply_load), then
  read the SoA back; ``load_to_device`` keeps it on the device;
* ``quantize_scene`` is QuantizedGaussianSceneSoa::from_soa_with_degree on the
  device (f16 pairs, octahedral rotations, trimmed f16 SH);
* ``PlyLoader.write_binary_ply_bytes`` is the canonical header plus the
  records of to_raw_vertices computed on the device (cgs_export_raw).
"""
from __future__ import annotations

import ctypes as C
import mmap
import os

import numpy as np

from . import _lib
from ._lib import PLY_ERRORS, GpuError, PlyLayout
from .pipeline import (DeviceBuffer, GaussianSceneSoa, GpuPipelineState,
                       QuantizedGaussianSceneSoa, _ptr, check)

# the 62 RawPlyVertex floats (raw_vertex.rs:16-37), in canonical property order
CANONICAL_PROPERTIES = (["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)]
                        + [f"f_rest_{i}" for i in range(45)] + ["opacity"]
                        + [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])


class PlyError(GpuError):
    """PlyError (ply_loader.rs:157-222): .ply_kind names the variant,
    .detail holds its numbers (count, or needed / available bytes)."""

    def __init__(self, code: int, message: str, kind: int, detail):
        super().__init__(code, message)
        self.ply_kind = PLY_ERRORS.get(kind, f"Ply{kind}")
        self.detail = tuple(int(d) for d in detail)


def _raise_ply(rc: int, layout: PlyLayout):
    if rc != _lib.CGS_OK:
        msg = _lib.load().cgs_last_error()
        raise PlyError(rc, msg.decode() if msg else "", layout.error, layout.error_detail)


def _bytes_ptr(data):
    """(pointer, length, keepalive) of a bytes-like object, zero-copy when possible."""
    if isinstance(data, np.ndarray):
        a = np.ascontiguousarray(data).view(np.uint8).ravel()
        return a.ctypes.data_as(C.c_void_p), a.nbytes, a
    mv = memoryview(data)
    if mv.readonly:
        a = np.frombuffer(mv, dtype=np.uint8)
        return a.ctypes.data_as(C.c_void_p), a.nbytes, a
    buf = (C.c_uint8 * mv.nbytes).from_buffer(mv)
    return C.cast(buf, C.c_void_p), mv.nbytes, buf


def parse_header(data) -> PlyLayout:
    """parse_header (ply_loader.rs:290-468) of a PLY file in memory."""
    p, n, keep = _bytes_ptr(data)
    layout = PlyLayout()
    _raise_ply(_lib.load().cgs_ply_parse_header(p, n, C.byref(layout)), layout)
    del keep
    return layout


class PlyLoader:
    """PlyLoader (ply_loader.rs:601-716) with the decode on the GPU."""

    @staticmethod
    def load_to_device(state: GpuPipelineState, data):
        """(count, {"pos", "cov", "opa", "sh"} device buffers) of a PLY in memory."""
        layout = parse_header(data)
        n = int(layout.vertex_count)
        p, ln, keep = _bytes_ptr(data)
        avail = max(0, ln - int(layout.payload_offset))
        if n * int(layout.vertex_stride) > avail:
            # checked before any buffer is sized by the (possibly huge) count;
            # the library applies the same test, without 64-bit wrap-around
            need = min(n * int(layout.vertex_stride), (1 << 64) - 1)
            layout.error = 13
            layout.error_detail[0] = need
            layout.error_detail[1] = avail
            raise PlyError(_lib.CGS_ERR_FORMAT,
                           f"vertex payload is truncated: need {need} bytes, file provides {avail}",
                           layout.error, layout.error_detail)
        cap = max(n, 1)
        out = {"pos": state.to_device(np.zeros((cap, 3), np.float32)),
               "cov": state.to_device(np.zeros((cap, 6), np.float32)),
               "opa": state.to_device(np.zeros(cap, np.float32)),
               "sh": state.to_device(np.zeros((cap, 48), np.float32))}
        rc = state._lib.cgs_ply_load(state._h, p, ln, C.byref(layout), out["pos"].ptr,
                                     out["cov"].ptr, out["opa"].ptr, out["sh"].ptr)
        del keep
        if rc != _lib.CGS_OK:
            for b in out.values():
                b.free()
            if rc == _lib.CGS_ERR_FORMAT:  # the header parsed: the payload is short
                layout.error = 13
                layout.error_detail[0] = n * layout.vertex_stride
                layout.error_detail[1] = max(0, ln - int(layout.payload_offset))
                _raise_ply(rc, layout)
            check(rc)
        return n, out

    @staticmethod
    def load_from_bytes(data, state: GpuPipelineState | None = None) -> GaussianSceneSoa:
        own = state is None
        st = GpuPipelineState(0) if own else state
        try:
            n, out = PlyLoader.load_to_device(st, data)
            scene = GaussianSceneSoa(out["pos"].numpy()[:n], out["cov"].numpy()[:n],
                                     out["opa"].numpy()[:n], out["sh"].numpy()[:n])
            for b in out.values():
                b.free()
            return scene
        finally:
            if own:
                st.close()

    @staticmethod
    def load_file(path, state: GpuPipelineState | None = None) -> GaussianSceneSoa:
        with open(path, "rb") as f:
            size = os.fstat(f.fileno()).st_size
            if size == 0:
                return PlyLoader.load_from_bytes(b"", state)
            with mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as m:
                return PlyLoader.load_from_bytes(np.frombuffer(m, dtype=np.uint8), state)

    @staticmethod
    def header_text(count: int) -> str:
        """The canonical header of write_binary_ply_bytes (ply_loader.rs:674-705)."""
        lines = ["ply", "format binary_little_endian 1.0", f"element vertex {count}"]
        lines += [f"property float {name}" for name in CANONICAL_PROPERTIES]
        return "\n".join(lines + ["end_header"]) + "\n"

    @staticmethod
    def write_binary_ply_bytes(scene: GaussianSceneSoa,
                               state: GpuPipelineState | None = None) -> bytes:
        own = state is None
        st = GpuPipelineState(0) if own else state
        try:
            n = scene.count
            body = b""
            if n:
                bufs = [st.to_device(a) for a in (scene.positions, scene.covariances_3d,
                                                   scene.opacities, scene.sh_coeffs)]
                rec = st.to_device(np.zeros((n, 62), np.float32))
                check(st._lib.cgs_export_raw(st._h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                             bufs[3].ptr, n, rec.ptr))
                body = rec.numpy().tobytes()
                for b in bufs + [rec]:
                    b.free()
            return PlyLoader.header_text(n).encode() + body
        finally:
            if own:
                st.close()

    @staticmethod
    def save_file(path, scene: GaussianSceneSoa, state: GpuPipelineState | None = None):
        with open(path, "wb") as f:
            f.write(PlyLoader.write_binary_ply_bytes(scene, state))


def sh_words(degree: int) -> int:
    """sh_packed_words_for_degree (soa.rs:342-345)."""
    bases = {0: 1, 1: 4, 2: 9}.get(int(degree), 16)
    return (bases * 3 + 1) // 2


def quantize_scene(state: GpuPipelineState, scene: GaussianSceneSoa,
                   sh_degree: int = 3) -> QuantizedGaussianSceneSoa:
    """QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:454-500) on the device."""
    n, deg = scene.count, min(int(sh_degree), 3)
    words = sh_words(deg)
    cov = np.zeros((n, 3), np.uint32)
    rot = np.zeros(n, np.uint32)
    shw = np.zeros((n, words), np.uint32)
    if n:
        dc, ds = state.to_device(scene.covariances_3d), state.to_device(scene.sh_coeffs)
        oc, orot, osh = (state.to_device(cov), state.to_device(rot), state.to_device(shw))
        check(state._lib.cgs_quantize_scene(state._h, dc.ptr, ds.ptr, n, deg, oc.ptr, orot.ptr,
                                            osh.ptr))
        cov, rot, shw = oc.numpy(), orot.numpy(), osh.numpy()
        for b in (dc, ds, oc, orot, osh):
            b.free()
    return QuantizedGaussianSceneSoa(scene.positions.copy(), cov, scene.opacities.copy(), shw,
                                     deg, rot)
