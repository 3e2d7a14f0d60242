"""Writes tests/golden/generator_fixtures.json: the first records of every
synthetic workload, as the seeded generator (cindergs_amd.rng and the
generators of cindergs_amd.synthetic / cages / mvc) draws them.  Floats are
stored as the bits of their float32 values.

    python tests/golden/make_generator_fixtures.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cindergs_amd import synthetic as syn  # noqa: E402
from cindergs_amd.cages import TetMesh  # noqa: E402
from cindergs_amd.mvc import icosphere  # noqa: E402

K = 64


def bits(a) -> list:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32).ravel().tolist()


def scene_rows(scene, lo, hi) -> dict:
    return {"first": lo, "count": hi - lo,
            "positions": bits(scene.positions[lo:hi]),
            "covariances": bits(scene.covariances_3d[lo:hi]),
            "opacities": bits(scene.opacities[lo:hi]),
            "sh": bits(scene.sh_coeffs[lo:hi])}


def main() -> None:
    out = {}
    ball = syn.scene_gaussian_ball()
    out["config1_ball_100k"] = scene_rows(ball, 0, K)
    cl = syn.scene_clusters()
    n_bg = int(round(cl.count * 0.10))
    n_cl = cl.count - n_bg
    out["config2_clusters_members"] = scene_rows(cl, 0, K)
    out["config2_clusters_background"] = scene_rows(cl, n_cl, n_cl + K)
    mesh = TetMesh.jittered_grid([-2.0, -1.5, -1.0], [2.0, 1.5, 1.0])
    out["config3_grid_vertices"] = {"lo": [-2.0, -1.5, -1.0], "hi": [2.0, 1.5, 1.0],
                                    "cells": [10, 10, 10], "jitter": 0.15, "seed": 3,
                                    "first": 0, "count": 2 * K,
                                    "vertices": bits(mesh.rest_vertices[:2 * K])}
    cage = icosphere()
    out["config4_icosphere_vertices"] = {"first": 0, "count": K,
                                         "vertices": bits(cage.vertices[:K])}
    vs, cams, comps = syn.batch_views(64, 3, seed=5)
    out["config5_views"] = {"count": 64, "scene": [int(v) for v in vs],
                            "size": [[c.screen_width, c.screen_height] for c in comps],
                            "eye": bits([list(c.camera_position) for c in comps])}
    path = os.path.join(ROOT, "tests", "golden", "generator_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
