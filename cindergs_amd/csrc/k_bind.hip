// k_bind.hip — splat-to-tetrahedron binding on the GPU.
//
// Same result as TetMesh::bind_point (src/deformation/mod.rs:267-310) run by
// bind_gaussians / bind_gaussians_with_bounds (263-323) for every splat: the
// first tetrahedron IN INDEX ORDER whose barycentric penalty
// sum max(-w_k, 0) is <= 1e-4, otherwise the tetrahedron of smallest penalty
// (first on ties); with bounds, a point outside them or whose best penalty
// exceeds 0.15 is unbound.  The weights are computed with the reference's
// operation order (w = inv_dm (p - X0) as glam's column sum, w0 = 1 - (w1 +
// w2 + w3)), so tet indices and weights are bit-exact.
//
// MI355X design: the reference scans all T tetrahedra per point (6000 x 1.07M
// tests for the cage workload).  Here a uniform grid over the cage holds, per
// cell, the tetrahedra whose acceptance region can reach the cell, in index
// order (built on the host, see runtime.cpp): every tetrahedron with penalty
// <= 1e-4 at p lies in p's cell list, so the first hit in that list is the
// first hit in index order.  Points with no hit in their cell (outside the
// cage, or in a gap) fall back to the full scan in index order, which is
// exactly the reference loop.  One thread per point; the per-tet record
// (X0, inv_dm) is 48 bytes.
#include "cgs_internal.h"

namespace {

// bind_point's weights and penalty (mod.rs:280-293); rec = X0[3], inv_dm[9]
// column-major
__device__ __forceinline__ float tet_weights(const float* rec, float px, float py, float pz,
                                             float w[4]) {
  const float dx = px - rec[0], dy = py - rec[1], dz = pz - rec[2];
  const float* m = rec + 3;
  w[1] = (m[0] * dx + m[3] * dy) + m[6] * dz;
  w[2] = (m[1] * dx + m[4] * dy) + m[7] * dz;
  w[3] = (m[2] * dx + m[5] * dy) + m[8] * dz;
  w[0] = 1.0f - ((w[1] + w[2]) + w[3]);
  return ((cgs_fmax(-w[0], 0.0f) + cgs_fmax(-w[1], 0.0f)) + cgs_fmax(-w[2], 0.0f)) +
         cgs_fmax(-w[3], 0.0f);
}

__global__ __launch_bounds__(256) void k_bind_tets(CgsBindArgs A) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= A.n) return;
  const float px = A.pos[(size_t)i * 3 + 0], py = A.pos[(size_t)i * 3 + 1],
              pz = A.pos[(size_t)i * 3 + 2];
  float* ow = A.out_w + (size_t)i * 4;
  if (A.has_bounds && (px < A.bmin[0] || px > A.bmax[0] || py < A.bmin[1] ||
                       py > A.bmax[1] || pz < A.bmin[2] || pz > A.bmax[2])) {
    A.out_tet[i] = 0xFFFFFFFFu;
    ow[0] = ow[1] = ow[2] = ow[3] = 0.0f;
    return;
  }
  float w[4];
  if (A.cell_start) {
    const float fx = (px - A.lo[0]) * A.inv_cell[0];
    const float fy = (py - A.lo[1]) * A.inv_cell[1];
    const float fz = (pz - A.lo[2]) * A.inv_cell[2];
    if (fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)A.dims[0] &&
        fy < (float)A.dims[1] && fz < (float)A.dims[2]) {
      const uint32_t c = ((uint32_t)fz * A.dims[1] + (uint32_t)fy) * A.dims[0] + (uint32_t)fx;
      const uint32_t j0 = A.cell_start[c], j1 = A.cell_start[c + 1];
      for (uint32_t j = j0; j < j1; ++j) {
        const uint32_t t = A.cell_tets[j];
        if (tet_weights(A.tet_rec + (size_t)t * 12, px, py, pz, w) <= 1e-4f) {
          A.out_tet[i] = t;
          ow[0] = w[0];
          ow[1] = w[1];
          ow[2] = w[2];
          ow[3] = w[3];
          return;
        }
      }
    }
  }
  // the reference loop: first acceptable tet, else the smallest penalty
  uint32_t best_t = 0;
  float best_pen = __builtin_inff();
  float bw[4] = {1.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t t = 0; t < A.num_tets; ++t) {
    const float pen = tet_weights(A.tet_rec + (size_t)t * 12, px, py, pz, w);
    if (pen <= 1e-4f) {
      A.out_tet[i] = t;
      ow[0] = w[0];
      ow[1] = w[1];
      ow[2] = w[2];
      ow[3] = w[3];
      return;
    }
    if (pen < best_pen) {
      best_pen = pen;
      best_t = t;
      bw[0] = w[0];
      bw[1] = w[1];
      bw[2] = w[2];
      bw[3] = w[3];
    }
  }
  if (A.has_bounds && best_pen > 0.15f) {
    A.out_tet[i] = 0xFFFFFFFFu;
    ow[0] = ow[1] = ow[2] = ow[3] = 0.0f;
    return;
  }
  A.out_tet[i] = best_t;
  ow[0] = bw[0];
  ow[1] = bw[1];
  ow[2] = bw[2];
  ow[3] = bw[3];
}

}  // namespace

void cgs_launch_bind_tets(const CgsBindArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_bind_tets, dim3((a.n + 255u) / 256u), dim3(256), 0, s, a);
}
