// cgs_scene_math.h — per-splat math of the MVC scene deformation.
//
// The operations of PolyCageMesh::deform_scene_cpu (src/deformation/mvc.rs:
// 564-749) after the per-point position and Jacobian: the polar rotation of
// F (higham_polar_decomposition, polar.rs:13-42), Sigma' = F Sigma F^T, the
// real-SH rotation of the coefficients (ShRotator, sh_rotation.rs:372-456),
// the strain test (jacobian_singular_values, polar.rs:89-109, through
// symmetric_eigen_3x3, math/covariance.rs:181-262), the major-axis split
// (split_gaussian_along_major_axis, 343-381) and the 3D mip filter
// (apply_mip_filter_3d, 288-304).
//
// 3x3 matrices are column-major float[9] (column k = m[3k..3k+2]) and every
// f32 expression keeps glam 0.28's operation order (Vec3 cross / dot, Mat3
// mul / inverse / determinant are scalar in glam for Mat3), so that with
// -ffp-contract=off the HIP kernel and the CPU oracle, which both include this
// file, produce the same bits.  The eigen solver and the SH rotation matrices
// are built in double as the reference builds them.
#pragma once

#include <stdint.h>

#include "cgs_detmath.h"

// ---- glam Vec3 / Mat3 (f32, column-major) ----

CGS_HD void cgs_v3_cross(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - b[1] * a[2];
  o[1] = a[2] * b[0] - b[2] * a[0];
  o[2] = a[0] * b[1] - b[0] * a[1];
}

CGS_HD float cgs_v3_dot(const float* a, const float* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Mat3::determinant: z_axis . (x_axis x y_axis)
CGS_HD float cgs_m3_det(const float* m) {
  float t[3];
  cgs_v3_cross(m, m + 3, t);
  return cgs_v3_dot(m + 6, t);
}

// Mat3::mul_mat3: column j of A B = (A.x b.x + A.y b.y) + A.z b.z, b = B col j
CGS_HD void cgs_m3_mul(const float* A, const float* B, float* o) {
  float r[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      r[3 * j + i] = A[i] * B[3 * j] + A[3 + i] * B[3 * j + 1] + A[6 + i] * B[3 * j + 2];
  for (int k = 0; k < 9; ++k) o[k] = r[k];
}

CGS_HD void cgs_m3_transpose(const float* m, float* o) {
  float r[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) r[3 * j + i] = m[3 * i + j];
  for (int k = 0; k < 9; ++k) o[k] = r[k];
}

// Mat3::inverse: cofactor columns y x z, z x x, x x y scaled by 1 / det, then
// transposed
CGS_HD void cgs_m3_inverse(const float* m, float* o) {
  float t0[3], t1[3], t2[3];
  cgs_v3_cross(m + 3, m + 6, t0);
  cgs_v3_cross(m + 6, m, t1);
  cgs_v3_cross(m, m + 3, t2);
  const float det = cgs_v3_dot(m + 6, t2);
  const float inv = 1.0f / det;
  const float* t[3] = {t0, t1, t2};
  // from_cols(t0 inv, t1 inv, t2 inv).transpose(): o[col j][row i] = t_i[j] inv
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) o[3 * j + i] = t[i][j] * inv;
}

// covariance_matrix_from_slice (math/covariance.rs:99-105)
CGS_HD void cgs_cov_to_m3(const float* c, float* m) {
  m[0] = c[0]; m[1] = c[1]; m[2] = c[2];
  m[3] = c[1]; m[4] = c[3]; m[5] = c[4];
  m[6] = c[2]; m[7] = c[4]; m[8] = c[5];
}

CGS_HD void cgs_m3_to_cov(const float* m, float* c) {
  c[0] = m[0]; c[1] = m[1]; c[2] = m[2];
  c[3] = m[4]; c[4] = m[5]; c[5] = m[8];
}

// ---- polar rotation (Higham, 5 Newton-Schulz steps) ----

// R of higham_polar_decomposition (polar.rs:13-42): identity for
// |det J| < 1e-8; R0 = J |det J|^(-1/3); R <- (R + R^-T) / 2 five times; the
// third column negated when det R < 0.
CGS_HD void cgs_polar_rotation(const float* J, float* R) {
  const float det = cgs_m3_det(J);
  if (fabsf(det) < 1e-8f) {
    for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    return;
  }
  const float gamma = cgs_powf_pos(fabsf(det), -1.0f / 3.0f);
  float r[9];
  for (int k = 0; k < 9; ++k) r[k] = J[k] * gamma;
  for (int it = 0; it < 5; ++it) {
    float ri[9], rit[9];
    cgs_m3_inverse(r, ri);
    cgs_m3_transpose(ri, rit);
    for (int k = 0; k < 9; ++k) r[k] = (r[k] + rit[k]) * 0.5f;
  }
  if (cgs_m3_det(r) < 0.0f)
    for (int k = 6; k < 9; ++k) r[k] = -r[k];
  for (int k = 0; k < 9; ++k) R[k] = r[k];
}

// ---- symmetric 3x3 eigen decomposition (cyclic Jacobi in double) ----

// symmetric_eigen_3x3 (math/covariance.rs:181-262): 12 sweeps of (0,1),
// (0,2), (1,2) rotations in double, eigenvalues sorted descending (stable),
// unit eigenvectors as the columns of a right-handed f32 basis.
CGS_HD void cgs_sym_eigen(const float* cov, float* evals, float* basis) {
  double a[3][3] = {{(double)cov[0], (double)cov[1], (double)cov[2]},
                    {(double)cov[1], (double)cov[3], (double)cov[4]},
                    {(double)cov[2], (double)cov[4], (double)cov[5]}};
  double v[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 12; ++sweep) {
    for (int pq = 0; pq < 3; ++pq) {
      const int p = P[pq], q = Q[pq];
      const double apq = a[p][q];
      if (fabs(apq) <= 1e-15) continue;
      const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                  : -1.0 / (-tau + sqrt(1.0 + tau * tau));
      const double c = 1.0 / sqrt(1.0 + t * t);
      const double s = t * c;
      const double app = a[p][p], aqq = a[q][q];
      a[p][p] = app - t * apq;
      a[q][q] = aqq + t * apq;
      a[p][q] = 0.0;
      a[q][p] = 0.0;
      for (int r = 0; r < 3; ++r) {
        if (r != p && r != q) {
          const double arp = a[r][p], arq = a[r][q];
          a[r][p] = c * arp - s * arq;
          a[p][r] = a[r][p];
          a[r][q] = s * arp + c * arq;
          a[q][r] = a[r][q];
        }
      }
      for (int r = 0; r < 3; ++r) {
        const double vrp = v[r][p], vrq = v[r][q];
        v[r][p] = c * vrp - s * vrq;
        v[r][q] = s * vrp + c * vrq;
      }
    }
  }
  // eigenpairs (value, normalised column); DVec3::normalize = v * (1 / |v|)
  double val[3], vec[3][3];
  for (int k = 0; k < 3; ++k) {
    val[k] = a[k][k];
    const double len = sqrt(v[0][k] * v[0][k] + v[1][k] * v[1][k] + v[2][k] * v[2][k]);
    const double rl = 1.0 / len;
    for (int r = 0; r < 3; ++r) vec[k][r] = v[r][k] * rl;
  }
  // stable descending sort of the three pairs
  int idx[3] = {0, 1, 2};
  for (int i = 1; i < 3; ++i) {
    const int key = idx[i];
    int j = i - 1;
    while (j >= 0 && val[idx[j]] < val[key]) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = key;
  }
  for (int c = 0; c < 3; ++c) {
    evals[c] = (float)val[idx[c]];
    for (int r = 0; r < 3; ++r) basis[3 * c + r] = (float)vec[idx[c]][r];
  }
  if (cgs_m3_det(basis) < 0.0f)
    for (int r = 0; r < 3; ++r) basis[6 + r] = -basis[6 + r];
}

// largest singular value of J (JacobianStrain::max_stretch): sqrt of the
// largest eigenvalue of J^T J
CGS_HD float cgs_max_stretch(const float* J) {
  float jt[9], c[9], cov[6], ev[3], basis[9];
  cgs_m3_transpose(J, jt);
  cgs_m3_mul(jt, J, c);
  cgs_m3_to_cov(c, cov);
  cgs_sym_eigen(cov, ev, basis);
  return cgs_sqrtf(cgs_fmax(ev[0], 0.0f));
}

// ---- 3D mip filter and major-axis split ----

// covariance_determinant (math/covariance.rs:113-125): in double, rounded
CGS_HD float cgs_cov_det64(const float* c) {
  const double xx = c[0], xy = c[1], xz = c[2], yy = c[3], yz = c[4], zz = c[5];
  const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - xz * yz) + xz * (xy * yz - xz * yy);
  return (float)det;
}

// apply_mip_filter_3d (288-304), in place
CGS_HD void cgs_mip_filter_3d(float* cov, float* alpha, float var) {
  if (var <= 0.0f) return;
  const float det_orig = cgs_fmax(cgs_cov_det64(cov), 1e-24f);
  float reg[6] = {cov[0] + var, cov[1], cov[2], cov[3] + var, cov[4], cov[5] + var};
  const float det_reg = cgs_fmax(cgs_cov_det64(reg), det_orig);
  const float comp = cgs_sqrtf(cgs_fclamp(det_orig / det_reg, 0.0f, 1.0f));
  for (int k = 0; k < 6; ++k) cov[k] = reg[k];
  *alpha = cgs_fclamp(*alpha * comp, 0.0f, 1.0f);
}

// split_gaussian_along_major_axis (343-381): the children share the covariance
// basis diag(max(l0/4, m), max(l1, m), max(l2, m)) basis^T (m = MIN_SCALE^2)
// and sit at -/+ offset along the unit major axis, |offset| = sqrt(max(l0, m)) / 2
CGS_HD void cgs_split_major(const float* cov, float* child_cov, float* offset) {
  float ev[3], basis[9];
  cgs_sym_eigen(cov, ev, basis);
  const float min_var = 1e-4f * 1e-4f;
  const float l0 = cgs_fmax(ev[0] * 0.25f, min_var);
  const float l1 = cgs_fmax(ev[1], min_var);
  const float l2 = cgs_fmax(ev[2], min_var);
  // Vec3::normalize_or_zero of the major axis
  const float* ax = basis;
  const float rcp = 1.0f / cgs_sqrtf(cgs_v3_dot(ax, ax));
  float major[3] = {0.0f, 0.0f, 0.0f};
  if (__builtin_isfinite(rcp) && rcp > 0.0f)
    for (int k = 0; k < 3; ++k) major[k] = ax[k] * rcp;
  const float mag = 0.5f * cgs_sqrtf(cgs_fmax(ev[0], min_var));
  for (int k = 0; k < 3; ++k) offset[k] = major[k] * mag;
  float diag[9] = {l0, 0.0f, 0.0f, 0.0f, l1, 0.0f, 0.0f, 0.0f, l2};
  float bt[9], bd[9], cm[9];
  cgs_m3_mul(basis, diag, bd);
  cgs_m3_transpose(basis, bt);
  cgs_m3_mul(bd, bt, cm);
  cgs_m3_to_cov(cm, child_cov);
}

// ---- real spherical-harmonics rotation, degrees 1-3 ----
//
// For band l the coefficient block c (2l+1 values, m = -l..l, the evaluate_sh
// order) rotates as c' = M c with M = (B T(R) B^+)^T (sh_rotation.rs:1-30):
// B [(2l+1) x n_l] writes the band's basis functions in the n_l = (l+1)(l+2)/2
// monomials of degree l, B^+ = B^T (B B^T)^-1 is its right pseudo-inverse and
// T(R) [n_l x n_l] maps the monomials v(d) to v(R^T d).  B and B^+ are
// constants (cgs_sh_basis_tables); T is expanded per rotation.

// monomial exponents of degree l in the reference's order (x descending,
// then y descending): exps_of (sh_rotation.rs:48-58)
CGS_HD int cgs_sh_nmono(int l) { return (l + 1) * (l + 2) / 2; }
CGS_HD void cgs_sh_exps(int l, int k, int* e) {
  e[0] = e[1] = e[2] = 0;
  int idx = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y) {
      if (idx == k) {
        e[0] = x;
        e[1] = y;
        e[2] = l - x - y;
        return;
      }
      ++idx;
    }
}
CGS_HD int cgs_sh_mono_index(int l, int x, int y) {
  // position of (x, y, l - x - y) in the order above
  int idx = 0;
  for (int xx = l; xx > x; --xx) idx += l - xx + 1;
  return idx + (l - x - y);
}

// Offsets of the per-degree tables in a packed double array: for l = 1, 2, 3
// the basis B (rows x n_l) then the pseudo-inverse (n_l x rows).
CGS_HD int cgs_sh_table_offset(int l, int which) {
  // sizes: l=1: 3x3 + 3x3; l=2: 5x6 + 6x5; l=3: 7x10 + 10x7
  const int sz1 = 9, sz2 = 30, sz3 = 70;
  int off = 0;
  if (l >= 2) off += 2 * sz1;
  if (l >= 3) off += 2 * sz2;
  const int sz = l == 1 ? sz1 : l == 2 ? sz2 : sz3;
  return off + (which ? sz : 0);
}
#define CGS_SH_TABLE_DOUBLES 218

// Row k of T(R): the monomial k of R^T d expanded in the monomials of d.  The
// linear forms are the columns of R (the rows of R^T); a monomial
// x^a y^b z^c is built by multiplying in a factors of form 0, then b of form
// 1, then c of form 2 (poly_mul_linear, sh_rotation.rs:281-305).
CGS_HD void cgs_sh_T_row(const float* R, int l, int k, double* out) {
  double forms[3][3];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) forms[a][c] = (double)R[3 * a + c];
  int e[3];
  cgs_sh_exps(l, k, e);
  double poly[10], next[10];
  int deg = 0;
  poly[0] = 1.0;
  for (int axis_form = 0; axis_form < 3; ++axis_form) {
    for (int rep = 0; rep < e[axis_form]; ++rep) {
      const int nn = cgs_sh_nmono(deg + 1);
      for (int j = 0; j < nn; ++j) next[j] = 0.0;
      const int na = cgs_sh_nmono(deg);
      for (int ia = 0; ia < na; ++ia) {
        const double coeff = poly[ia];
        if (coeff == 0.0) continue;
        int pe[3];
        cgs_sh_exps(deg, ia, pe);
        for (int ax = 0; ax < 3; ++ax) {
          const double f = forms[axis_form][ax];
          if (f == 0.0) continue;
          int ne[3] = {pe[0], pe[1], pe[2]};
          ne[ax] += 1;
          next[cgs_sh_mono_index(deg + 1, ne[0], ne[1])] += coeff * f;
        }
      }
      ++deg;
      for (int j = 0; j < nn; ++j) poly[j] = next[j];
    }
  }
  const int n = cgs_sh_nmono(l);
  for (int j = 0; j < n; ++j) out[j] = poly[j];
}

// The (2l+1)^2 row-major f32 rotation matrix of band l (sh_rotation_matrix,
// sh_rotation.rs:312-333): C = B T, D = C B^+, M = D^T; the products skip zero
// left factors and sum their inner index in ascending order, as mat_mul does.
CGS_HD void cgs_sh_band_matrix(const float* R, int l, const double* tables, float* M) {
  const int rows = 2 * l + 1, n = cgs_sh_nmono(l);
  const double* B = tables + cgs_sh_table_offset(l, 0);     // rows x n
  const double* Pinv = tables + cgs_sh_table_offset(l, 1);  // n x rows
  double C[7][10];
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < n; ++j) C[i][j] = 0.0;
  for (int k = 0; k < n; ++k) {
    double t[10];
    cgs_sh_T_row(R, l, k, t);
    for (int i = 0; i < rows; ++i) {
      const double bik = B[i * n + k];
      if (bik == 0.0) continue;
      for (int j = 0; j < n; ++j) C[i][j] += bik * t[j];
    }
  }
  for (int i = 0; i < rows; ++i) {
    double d[7];
    for (int j = 0; j < rows; ++j) d[j] = 0.0;
    for (int k = 0; k < n; ++k) {
      const double cik = C[i][k];
      if (cik == 0.0) continue;
      for (int j = 0; j < rows; ++j) d[j] += cik * Pinv[k * rows + j];
    }
    // M = D^T, row-major: M[j][i] = D[i][j]
    for (int j = 0; j < rows; ++j) M[j * rows + i] = (float)d[j];
  }
}

// ShRotator::apply (sh_rotation.rs:420-438) for one band of one channel:
// c'[r] = sum_c M[r][c] c[c], accumulated in f32 in column order
CGS_HD void cgs_sh_apply_block(float* coeffs, int off, int size, const float* M) {
  float rot[7];
  for (int r = 0; r < size; ++r) {
    float acc = 0.0f;
    for (int c = 0; c < size; ++c) acc += M[r * size + c] * coeffs[off + c];
    rot[r] = acc;
  }
  for (int r = 0; r < size; ++r) coeffs[off + r] = rot[r];
}

// rotates the 48-coefficient record [dc r g b, rest_r 15, rest_g 15, rest_b 15]
// for degree 1..3 (degree 0: unchanged)
CGS_HD void cgs_sh_rotate(float* sh, const float* R, int degree, const double* tables) {
  if (degree <= 0) return;
  if (degree > 3) degree = 3;
  float M1[9], M2[25], M3[49];
  cgs_sh_band_matrix(R, 1, tables, M1);
  if (degree >= 2) cgs_sh_band_matrix(R, 2, tables, M2);
  if (degree >= 3) cgs_sh_band_matrix(R, 3, tables, M3);
  for (int ch = 0; ch < 3; ++ch) {
    const int rest = 3 + ch * 15;
    cgs_sh_apply_block(sh, rest, 3, M1);
    if (degree >= 2) cgs_sh_apply_block(sh, rest + 3, 5, M2);
    if (degree >= 3) cgs_sh_apply_block(sh, rest + 8, 7, M3);
  }
}

// ---- the constant tables (host) ----
//
// Basis rows of the real SH bands in the monomial basis (basis_rows,
// sh_rotation.rs:123-190; the evaluate_sh constants), the Gram matrix B B^T,
// its Gauss-Jordan inverse with partial pivoting (invert, 232-270) and the
// pseudo-inverse B^T (B B^T)^-1, packed by cgs_sh_table_offset.
static inline void cgs_sh_basis_tables(double* tables) {
  const double C1 = 0.4886025119029199, C20 = 1.0925484305920792, C22 = 0.31539156525252005,
               C24 = 0.5462742152960396, C30 = -0.5900435899266435, C31 = 2.890611442640554,
               C32 = -0.4570457994644658, C33 = 0.3731763325901154, C35 = 1.445305721320277;
  for (int i = 0; i < CGS_SH_TABLE_DOUBLES; ++i) tables[i] = 0.0;
  for (int l = 1; l <= 3; ++l) {
    const int rows = 2 * l + 1, n = cgs_sh_nmono(l);
    double* B = tables + cgs_sh_table_offset(l, 0);
    double* Pinv = tables + cgs_sh_table_offset(l, 1);
    auto set = [&](int row, int x, int y, double v) { B[row * n + cgs_sh_mono_index(l, x, y)] = v; };
    if (l == 1) {
      set(0, 0, 1, -C1);  // -C1 y
      set(1, 0, 0, C1);   //  C1 z
      set(2, 1, 0, -C1);  // -C1 x
    } else if (l == 2) {
      set(0, 1, 1, C20);         // C20 xy
      set(1, 0, 1, -C20);        // -C20 yz
      set(2, 0, 0, 2.0 * C22);   // C22 (2 zz - xx - yy)
      set(2, 2, 0, -1.0 * C22);
      set(2, 0, 2, -1.0 * C22);
      set(3, 1, 0, -C20);        // -C20 xz
      set(4, 2, 0, C24);         // C24 (xx - yy)
      set(4, 0, 2, -1.0 * C24);
    } else {
      set(0, 2, 1, 3.0 * C30);   // C30 y (3 xx - yy)
      set(0, 0, 3, -1.0 * C30);
      set(1, 1, 1, C31);         // C31 xyz
      set(2, 0, 1, 4.0 * C32);   // C32 y (4 zz - xx - yy)
      set(2, 2, 1, -1.0 * C32);
      set(2, 0, 3, -1.0 * C32);
      set(3, 0, 0, 2.0 * C33);   // C33 z (2 zz - 3 xx - 3 yy)
      set(3, 2, 0, -3.0 * C33);
      set(3, 0, 2, -3.0 * C33);
      set(4, 1, 0, 4.0 * C32);   // C32 x (4 zz - xx - yy)
      set(4, 3, 0, -1.0 * C32);
      set(4, 1, 2, -1.0 * C32);
      set(5, 2, 0, C35);         // C35 z (xx - yy)
      set(5, 0, 2, -1.0 * C35);
      set(6, 3, 0, C30);         // C30 x (xx - 3 yy)
      set(6, 1, 2, -3.0 * C30);
    }
    // G = B B^T, inverted by Gauss-Jordan with partial pivoting
    double aug[7][14];
    for (int i = 0; i < rows; ++i) {
      for (int j = 0; j < rows; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
          const double bik = B[i * n + k];
          if (bik == 0.0) continue;
          s += bik * B[j * n + k];
        }
        aug[i][j] = s;
        aug[i][rows + j] = i == j ? 1.0 : 0.0;
      }
    }
    for (int col = 0; col < rows; ++col) {
      // the row of largest |value| in the column, the last one on ties (the
      // element Iterator::max_by returns)
      int piv = col;
      for (int r = col + 1; r < rows; ++r)
        if (fabs(aug[r][col]) >= fabs(aug[piv][col])) piv = r;
      if (piv != col)
        for (int j = 0; j < 2 * rows; ++j) {
          const double t = aug[col][j];
          aug[col][j] = aug[piv][j];
          aug[piv][j] = t;
        }
      const double ip = 1.0 / aug[col][col];
      for (int j = 0; j < 2 * rows; ++j) aug[col][j] *= ip;
      for (int r = 0; r < rows; ++r) {
        if (r == col) continue;
        const double f = aug[r][col];
        if (f == 0.0) continue;
        for (int j = 0; j < 2 * rows; ++j) aug[r][j] -= f * aug[col][j];
      }
    }
    // B^+ = B^T G^-1 (n x rows)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < rows; ++j) {
        double s = 0.0;
        for (int i = 0; i < rows; ++i) {
          const double bik = B[i * n + k];
          if (bik == 0.0) continue;
          s += bik * aug[i][rows + j];
        }
        Pinv[k * rows + j] = s;
      }
  }
}

// The same rotation as straight-line code (generated by gen_sh_rotation.py):
// cgs_sh_rotate_fast(sh, R, degree, tables) gives the bits of cgs_sh_rotate.
#include "cgs_sh_rotation_gen.h"
