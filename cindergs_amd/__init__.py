"""cindergs_amd — MI355X (gfx950) implementation of CinderGS's splat render and
cage-deformation hot path behind a C ABI (include/cindergs.h).

The compute path is the HIP library libcindergs_hip.so built in-tree; this
package is its host-side mirror of the reference API (src/render/gpu_state.rs).
"""
from .pipeline import (CameraUniforms, CompositorUniforms, DeviceBuffer, FrameParams,  # noqa: F401
                       GaussianSceneSoa, GpuError, GpuPipelineState, QuantizedGaussianSceneSoa,
                       StageTimings, TetCage, ViewBatch, padded_rgba8_row_u32)
