"""The MVC surrogate network on the GPU (cgs_mlp_* entry points).

A coordinate MLP (Fourier features -> 4 x 256 ReLU -> n_out) used to bind
many points to one closed cage, with exact MVC as the fallback for any point
whose reconstruction residual |sum_j w_j v_j - x| is above a threshold (the
reference has no learned component: it always evaluates exact MVC,
src/deformation/mvc.rs:498-541).  Either the network stands for the whole
weights (affine-normalised so they sum to 1, trained on exact MVC weights), or,
with a near field, for the far share of MVC's sums only (trained on
MvcGpu.far_share targets) while the near sums are evaluated exactly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .mvc import PolyCage
from .pipeline import DeviceBuffer, GpuPipelineState, _ptr, check

LAYERS = 5


class MvcSurrogate:
    def __init__(self, state: GpuPipelineState, n_out: int, fourier_sigma: float = 2.0,
                 seed: int = 0xC1DE5EED00000004):
        self.state = state
        self._lib = state._lib
        h = C.c_void_p()
        check(self._lib.cgs_mlp_create(state._h, n_out, fourier_sigma, seed, C.byref(h)))
        self._h = h
        self.n_out = n_out
        self.near_field = (0.0, 0.0)
        self.far_kind = 0
        self.count = int(self._lib.cgs_mlp_param_count(h))
        rows = (C.c_uint32 * LAYERS)()
        cols = (C.c_uint32 * LAYERS)()
        offs = (C.c_uint64 * LAYERS)()
        check(self._lib.cgs_mlp_layout(h, rows, cols, offs))
        self.rows, self.cols, self.offsets = list(rows), list(cols), list(offs)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cgs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters -------------------------------------------------------
    def get_params(self):
        p = np.zeros(self.count, np.float32)
        f = np.zeros((32, 3), np.float32)
        check(self._lib.cgs_mlp_get_params(self._h, _ptr(p), _ptr(f)))
        return p, f

    def set_params(self, params=None, fourier=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        f = None if fourier is None else np.ascontiguousarray(fourier, np.float32)
        check(self._lib.cgs_mlp_set_params(self._h, _ptr(p), _ptr(f)))

    def layer(self, params: np.ndarray, l: int):
        """(W [rows x cols], b [rows]) of layer l from a flat parameter vector."""
        r, c, o = self.rows[l], self.cols[l], self.offsets[l]
        blk = params[o:o + r * (c + 32)].reshape(r, c + 32)
        return blk[:, :c], blk[:, c]

    def set_near_field(self, r0: float, r1: float, kind: int = 0):
        """Near/far split (cgs_mlp_set_near_field): the network then learns the
        far share of the weights (kind 0) or the far sums (kind 1,
        cgs_mlp_set_far_kind), with targets from MvcGpu.far_share(..., kind),
        and the binding adds the exactly evaluated near sums.  (0, 0) switches
        it off."""
        check(self._lib.cgs_mlp_set_near_field(self._h, float(r0), float(r1)))
        check(self._lib.cgs_mlp_set_far_kind(self._h, int(kind)))
        self.near_field = (float(r0), float(r1))
        self.far_kind = int(kind)

    # ---- training and inference ------------------------------------------
    def train(self, d_x: DeviceBuffer, d_t: DeviceBuffer, n: int, steps: int, batch: int = 8192,
              lr: float = 1e-3, record_loss: bool = True):
        losses = np.zeros(max(steps, 1), np.float32) if record_loss else None
        check(self._lib.cgs_mlp_train(self._h, d_x.ptr, d_t.ptr, n, steps, batch, lr,
                                      _ptr(losses)))
        return losses[:steps] if record_loss else None

    def gradient(self, d_x: DeviceBuffer, d_t: DeviceBuffer, n: int):
        g = np.zeros(self.count, np.float32)
        loss = np.zeros(1, np.float32)
        check(self._lib.cgs_mlp_gradient(self._h, d_x.ptr, d_t.ptr, n, _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def infer(self, d_x: DeviceBuffer, n: int, gradients: bool = True):
        w = self.state.to_device(np.zeros((n, self.n_out), np.float32))
        g = self.state.to_device(np.zeros((n, self.n_out, 3), np.float32)) if gradients else None
        check(self._lib.cgs_mlp_infer(self._h, d_x.ptr, n, w.ptr, g.ptr if g else None))
        return w, g

    def bind(self, cage: PolyCage, d_points: DeviceBuffer, n: int, residual_tol: float,
             gradients: bool = True):
        """Surrogate bind_gaussians: (weights, gradients, inside) device buffers
        and the number of points that took the exact-MVC fallback."""
        nv = cage.num_vertices
        w = self.state.to_device(np.zeros((n, nv), np.float32))
        g = self.state.to_device(np.zeros((n, nv, 3), np.float32)) if gradients else None
        m = self.state.to_device(np.zeros(n, np.uint32))
        fb = C.c_uint64()
        check(self._lib.cgs_mlp_bind(self.state._h, self._h, _ptr(cage.vertices), nv,
                                     _ptr(cage.triangles), cage.triangles.shape[0], d_points.ptr,
                                     n, residual_tol, w.ptr, g.ptr if g else None, m.ptr,
                                     C.byref(fb)))
        return w, g, m, int(fb.value)
