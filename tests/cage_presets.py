"""The reference's preset cages (create_box_surface_cage,
create_octahedron_cage, create_icosahedron_cage: src/deformation/mvc.rs:
113-243), built from the vertex formulas and triangle tables held in
golden/reference_fixtures.json."""
import json
import math
import os

import numpy as np

from cindergs_amd.mvc import PolyCage

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


def preset(kind, a, b):
    """("box", lo, hi), ("octahedron", centre, radius) or ("icosahedron",
    centre, radius)."""
    m = FIX["mvc_cages"]
    if kind == "box":
        lo, hi = a, b
        V = [[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]],
             [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
             [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]]
        return PolyCage(np.array(V, np.float32), np.array(m["box_triangles"]))
    c, r = np.array(a, np.float32), np.float32(max(abs(b), 1e-4))
    if kind == "octahedron":
        d = [[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r], [0, 0, -r]]
        return PolyCage(c + np.array(d, np.float32), np.array(m["octahedron_triangles"]))
    phi = (1.0 + math.sqrt(5.0)) * 0.5
    raw = np.array([[-1, phi, 0], [1, phi, 0], [-1, -phi, 0], [1, -phi, 0], [0, -1, phi],
                    [0, 1, phi], [0, -1, -phi], [0, 1, -phi], [phi, 0, -1], [phi, 0, 1],
                    [-phi, 0, -1], [-phi, 0, 1]], np.float32)
    raw = raw / np.linalg.norm(raw, axis=1, keepdims=True)
    return PolyCage(c + raw * r, np.array(m["icosahedron_triangles"]))


def scene_case(case):
    """(cage, deformed vertices, rest pos [n,3], cov [n,6], opacity [n], sh [n,48])
    of a FIX["mvc_scene"] entry."""
    if "octahedron" in case:
        cage = preset("octahedron", *case["octahedron"])
    else:
        cage = preset("box", *case["box"])
    Vd = cage.vertices.copy()
    for vi, ax, dv in case.get("deform", []):
        Vd[vi, ax] += np.float32(dv)
    if "translate" in case:
        Vd = (Vd + np.array(case["translate"], np.float32)).astype(np.float32)
    sp = case["splats"]
    pos = np.array([s["pos"] for s in sp], np.float32)
    cov = np.array([s["cov"] for s in sp], np.float32)
    opa = np.array([s["opacity"] for s in sp], np.float32)
    sh = np.array([[s["sh_fill"]] * 48 for s in sp], np.float32)
    return cage, Vd, pos, cov, opa, sh
