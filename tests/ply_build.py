"""PLY files for the ingest tests, built as the reference's tests build them
(header_text / canonical_header / the payload written property by property,
src/scene/ply_loader.rs:718-745, 777-803), and the slot table that
parse_vertex_payload (540-598) implies for a property list."""
import struct

import numpy as np

TYPES = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h",
         "ushort": "H", "uint16": "H", "int": "i", "int32": "i", "uint": "I", "uint32": "I",
         "float": "f", "float32": "f", "double": "d", "float64": "d"}
TYPE_IDS = {"b": 1, "B": 2, "h": 3, "H": 4, "i": 5, "I": 6, "f": 7, "d": 8}

STANDARD = [("float", n) for n in ("x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                   "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
                                   "rot_3")]
CANONICAL_NAMES = (["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)]
                   + [f"f_rest_{i}" for i in range(45)] + ["opacity"]
                   + [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])


def header_text(fmt, count, props, eol="\n"):
    lines = ["ply", f"format {fmt}", f"element vertex {count}"]
    lines += [f"property {t} {n}" for t, n in props]
    return eol.join(lines + ["end_header"]) + eol


def canonical_props():
    return [("float", n) for n in CANONICAL_NAMES]


def payload(props, rows, big_endian=False):
    """rows: [n][len(props)] values, written in declaration order."""
    fmt = (">" if big_endian else "<") + "".join(TYPES[t] for t, _ in props)
    return b"".join(struct.pack(fmt, *row) for row in rows)


def ply_bytes(props, rows, big_endian=False, fmt=None, eol="\n"):
    f = fmt or ("binary_big_endian 1.0" if big_endian else "binary_little_endian 1.0")
    return header_text(f, len(rows), props, eol).encode() + payload(props, rows, big_endian)


def slot_table(props):
    """(offsets [62], types [62], stride) that parse_vertex_payload reads: the
    first property of each name, f_rest scattered 15 per channel."""
    offs, first = {}, {}
    o = 0
    for t, n in props:
        c = TYPES[t]
        if n not in first:
            first[n] = (o, TYPE_IDS[c])
        o += struct.calcsize(c)
    so, st = np.zeros(62, np.uint32), np.zeros(62, np.uint8)

    def put(slot, name):
        if name in first:
            so[slot], st[slot] = first[name]

    for k, n in enumerate(["x", "y", "z", "nx", "ny", "nz"]):
        put(k, n)
    for c in range(3):
        put(6 + c, f"f_dc_{c}")
    rest = 0
    while f"f_rest_{rest}" in first:
        rest += 1
    per = rest // 3
    for c in range(3):
        for i in range(per):
            put(9 + 15 * c + i, f"f_rest_{c * per + i}")
    put(54, "opacity")
    for c in range(3):
        put(55 + c, f"scale_{c}")
    for c in range(4):
        put(58 + c, f"rot_{c}")
    del offs
    return so, st, o


def random_rows(props, n, seed):
    """Plausible 3DGS values per property (ints for integer properties)."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        row = []
        for t, name in props:
            c = TYPES[t]
            if c in "bBhHiI":
                lo, hi = {"b": (-128, 127), "B": (0, 255), "h": (-3000, 3000), "H": (0, 60000),
                          "i": (-100000, 100000), "I": (0, 4000000000)}[c]
                row.append(int(rng.integers(lo, hi)))
            elif name.startswith("scale"):
                row.append(float(rng.normal(-3.0, 1.0)))
            elif name.startswith("rot"):
                row.append(float(rng.normal()))
            elif name == "opacity":
                row.append(float(rng.normal(0.0, 2.0)))
            elif name in ("x", "y", "z"):
                row.append(float(rng.uniform(-5, 5)))
            else:
                row.append(float(rng.normal(0.0, 0.5)))
        rows.append(row)
    return rows
