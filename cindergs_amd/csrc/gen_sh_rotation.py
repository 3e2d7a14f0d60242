#!/usr/bin/env python3
"""Writes cgs_sh_rotation_gen.h: straight-line code for the real-SH band
rotation of cgs_scene_math.h (cgs_sh_band_matrix + cgs_sh_apply_block), with
the same floating-point operations in the same order.

For band l the rotated coefficients are c' = M c, M = (B T(R) B^+)^T
(sh_rotation.rs:1-333 of the reference).  The loop form in cgs_scene_math.h
builds the T rows by multiplying linear forms monomial by monomial, C = B T
skipping zero entries of B, D = C B^+ skipping zero entries of C, M = D^T,
and then accumulates c'[r] = sum_c M[r][c] c[c] in f32 in column order.
The generated form:
  * expands every T row at generation time (exponents and monomial indices
    are constants), keeping each `+=` as an explicit add of the same terms in
    the same order (a zero factor adds +0.0, which leaves an accumulator that
    started at +0.0 unchanged, as the skip does);
  * skips the zeros of B and writes the entries of B and B^+ as exact
    literals, computed here by the operations of cgs_sh_basis_tables;
  * computes C and D one row at a time, taking the rows of B in groups that
    share T rows (so that a T row lives only while its group is built), and
    then accumulates c' in f32 in the same column order.

Run from this directory: python3 gen_sh_rotation.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def nmono(l):
    return (l + 1) * (l + 2) // 2


def exps(l, k):
    idx = 0
    for x in range(l, -1, -1):
        for y in range(l - x, -1, -1):
            if idx == k:
                return (x, y, l - x - y)
            idx += 1
    raise ValueError


def mono_index(l, x, y):
    idx = 0
    for xx in range(l, x, -1):
        idx += l - xx + 1
    return idx + (l - x - y)


C1, C20, C22 = 0.4886025119029199, 1.0925484305920792, 0.31539156525252005
C24, C30, C31 = 0.5462742152960396, -0.5900435899266435, 2.890611442640554
C32, C33, C35 = -0.4570457994644658, 0.3731763325901154, 1.445305721320277

# basis rows as (row, x, y, value): the evaluate_sh polynomials of each band
B_SET = {
    1: [(0, 0, 1, -C1), (1, 0, 0, C1), (2, 1, 0, -C1)],
    2: [(0, 1, 1, C20), (1, 0, 1, -C20), (2, 0, 0, 2.0 * C22), (2, 2, 0, -1.0 * C22),
        (2, 0, 2, -1.0 * C22), (3, 1, 0, -C20), (4, 2, 0, C24), (4, 0, 2, -1.0 * C24)],
    3: [(0, 2, 1, 3.0 * C30), (0, 0, 3, -1.0 * C30), (1, 1, 1, C31), (2, 0, 1, 4.0 * C32),
        (2, 2, 1, -1.0 * C32), (2, 0, 3, -1.0 * C32), (3, 0, 0, 2.0 * C33), (3, 2, 0, -3.0 * C33),
        (3, 0, 2, -3.0 * C33), (4, 1, 0, 4.0 * C32), (4, 3, 0, -1.0 * C32), (4, 1, 2, -1.0 * C32),
        (5, 2, 0, C35), (5, 0, 2, -1.0 * C35), (6, 3, 0, C30), (6, 1, 2, -3.0 * C30)],
}


# non-zero monomials (x, y) of each basis row
B_ROWS = {l: [[(x, y) for (row, x, y, _) in B_SET[l] if row == i] for i in range(2 * l + 1)]
          for l in (1, 2, 3)}


def basis_tables():
    """The 218 doubles of cgs_sh_basis_tables, by the same operations in the
    same order (Python floats are IEEE binary64 with no contraction)."""
    t = [0.0] * 218
    for l in (1, 2, 3):
        rows, n = 2 * l + 1, nmono(l)
        ob, op = table_offset(l, 0), table_offset(l, 1)
        B = [0.0] * (rows * n)
        for row, x, y, v in B_SET[l]:
            B[row * n + mono_index(l, x, y)] = v
        aug = [[0.0] * (2 * rows) for _ in range(rows)]
        for i in range(rows):
            for j in range(rows):
                acc = 0.0
                for k in range(n):
                    if B[i * n + k] == 0.0:
                        continue
                    acc += B[i * n + k] * B[j * n + k]
                aug[i][j] = acc
                aug[i][rows + j] = 1.0 if i == j else 0.0
        for col in range(rows):
            piv = col
            for r in range(col + 1, rows):
                if abs(aug[r][col]) >= abs(aug[piv][col]):
                    piv = r
            if piv != col:
                aug[col], aug[piv] = aug[piv], aug[col]
            ip = 1.0 / aug[col][col]
            for j in range(2 * rows):
                aug[col][j] *= ip
            for r in range(rows):
                if r == col:
                    continue
                f = aug[r][col]
                if f == 0.0:
                    continue
                for j in range(2 * rows):
                    aug[r][j] -= f * aug[col][j]
        P = [0.0] * (n * rows)
        for k in range(n):
            for j in range(rows):
                acc = 0.0
                for i in range(rows):
                    if B[i * n + k] == 0.0:
                        continue
                    acc += B[i * n + k] * aug[i][rows + j]
                P[k * rows + j] = acc
        t[ob:ob + rows * n] = B
        t[op:op + n * rows] = P
    return t


TABLE = None


def lit(v):
    """exact C++ literal of a double"""
    if v == 0.0:
        return "0.0" if str(v)[0] != "-" else "-0.0"
    return v.hex()


def table_offset(l, which):
    sz = {1: 9, 2: 30, 3: 70}
    off = 0
    if l >= 2:
        off += 2 * sz[1]
    if l >= 3:
        off += 2 * sz[2]
    return off + (sz[l] if which else 0)


class Emitter:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self, expr, ty="double"):
        name = f"t{self.n}"
        self.n += 1
        self.lines.append(f"  const {ty} {name} = {expr};")
        return name


def t_row(em, l, k):
    """T row k of band l as a list of n expressions (variable names)."""
    e = exps(l, k)
    poly = ["1.0"]  # degree-0 polynomial
    deg = 0
    for form in range(3):
        for _ in range(e[form]):
            nn = nmono(deg + 1)
            nxt = ["0.0"] * nn
            for ia in range(nmono(deg)):
                coeff = poly[ia]
                pe = exps(deg, ia)
                for ax in range(3):
                    f = f"fm{form}{ax}"
                    ne = list(pe)
                    ne[ax] += 1
                    tgt = mono_index(deg + 1, ne[0], ne[1])
                    if coeff == "1.0":
                        term = f"({f} == 0.0 ? 0.0 : {f})"
                    else:
                        term = f"(({coeff} == 0.0 || {f} == 0.0) ? 0.0 : {coeff} * {f})"
                    nxt[tgt] = em.tmp(f"{nxt[tgt]} + {term}")
            deg += 1
            poly = nxt
    if deg == 0:
        poly = [em.tmp("1.0")]
    return poly


def groups(l):
    """Rows of B joined when they use a common monomial (T row), in order of
    their first row: each group's T rows are needed by that group only."""
    rows = len(B_ROWS[l])
    mono = [set(mono_index(l, x, y) for (x, y) in B_ROWS[l][i]) for i in range(rows)]
    parent = list(range(rows))

    def find(a):
        while parent[a] != a:
            a = parent[a]
        return a

    for i in range(rows):
        for j in range(i + 1, rows):
            if mono[i] & mono[j]:
                parent[find(j)] = find(i)
    out = {}
    for i in range(rows):
        out.setdefault(find(i), []).append(i)
    return sorted(out.values(), key=lambda g: g[0])


def band(em, l, matrix_only=False):
    """Band l's rotation applied to sh (or, with matrix_only, its matrix stored
    row-major into M, as cgs_sh_band_matrix stores it)."""
    rows, n = 2 * l + 1, nmono(l)
    ob, op = table_offset(l, 0), table_offset(l, 1)
    em.lines.append(f"  // band {l}: rest slots {l * l - 1}..{l * l + 2 * l - 1} of each channel")
    base = l * l - 1
    m = {}
    for grp in groups(l):
        em.lines.append(f"  // rows {grp} of B")
        trows = {}
        for i in grp:
            for (x, y) in B_ROWS[l][i]:
                k = mono_index(l, x, y)
                if k not in trows:
                    trows[k] = None
        for k in sorted(trows):
            trows[k] = t_row(em, l, k)
        for i in grp:
            # C row i = sum over the non-zero B[i][k], k ascending, of B[i][k] T[k]
            nz = sorted(mono_index(l, x, y) for (x, y) in B_ROWS[l][i])
            c = ["0.0"] * n
            for k in nz:
                b = lit(TABLE[ob + i * n + k])
                for j in range(n):
                    c[j] = em.tmp(f"{c[j]} + {b} * {trows[k][j]}")
            # D row i = sum over k ascending of C[i][k] B^+[k] (zero C entries add +0)
            d = ["0.0"] * rows
            for k in range(n):
                for j in range(rows):
                    pk = lit(TABLE[op + k * rows + j])
                    d[j] = em.tmp(f"{d[j]} + ({c[k]} == 0.0 ? 0.0 : {c[k]} * {pk})")
            # M[j][i] = (float) D[i][j]
            for j in range(rows):
                m[j, i] = em.tmp(f"(float){d[j]}", "float")
    if matrix_only:
        for j in range(rows):
            for i in range(rows):
                em.lines.append(f"  M[{j * rows + i}] = {m[j, i]};")
        return
    # c'[j] = sum over i ascending of M[j][i] c[i], per channel, in f32
    for ch in range(3):
        orig = [em.tmp(f"sh[{3 + ch * 15 + base + i}]", "float") for i in range(rows)]
        for j in range(rows):
            acc = "0.0f"
            for i in range(rows):
                acc = em.tmp(f"{acc} + {m[j, i]} * {orig[i]}", "float")
            em.lines.append(f"  const float r{ch}_{j} = {acc};")
        for j in range(rows):
            em.lines.append(f"  sh[{3 + ch * 15 + base + j}] = r{ch}_{j};")


def main():
    global TABLE
    TABLE = basis_tables()
    out = ["// cgs_sh_rotation_gen.h — generated by gen_sh_rotation.py; do not edit.",
           "//",
           "// Straight-line real-SH band rotation: the operations of cgs_sh_band_matrix",
           "// and cgs_sh_apply_block (cgs_scene_math.h) in the same order, for bands",
           "// 1-3 of the 48-coefficient record, with the entries of B and B^+",
           "// (cgs_sh_basis_tables) as exact literals.",
           "#pragma once", ""]
    for l in (1, 2, 3):
        em = Emitter()
        band(em, l)
        out.append(f"CGS_HD void cgs_sh_rotate_band{l}(float* sh, const float* R) {{")
        for a in range(3):
            for c in range(3):
                out.append(f"  const double fm{a}{c} = (double)R[{3 * a + c}];")
        out += em.lines
        out += ["}", ""]
    for l in (1, 2, 3):
        em = Emitter()
        band(em, l, matrix_only=True)
        out.append(f"// cgs_sh_band_matrix(R, {l}, tables, M) with the same operations")
        out.append(f"CGS_HD void cgs_sh_band{l}_matrix(const float* R, float* M) {{")
        for a in range(3):
            for c in range(3):
                out.append(f"  const double fm{a}{c} = (double)R[{3 * a + c}];")
        out += em.lines
        out += ["}", ""]
    out += ["// cgs_sh_rotate with the generated bands",
            "CGS_HD void cgs_sh_rotate_fast(float* sh, const float* R, int degree) {",
            "  if (degree <= 0) return;",
            "  cgs_sh_rotate_band1(sh, R);",
            "  if (degree >= 2) cgs_sh_rotate_band2(sh, R);",
            "  if (degree >= 3) cgs_sh_rotate_band3(sh, R);",
            "}", ""]
    with open(os.path.join(HERE, "cgs_sh_rotation_gen.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
