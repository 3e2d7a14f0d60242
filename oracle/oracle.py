"""ctypes wrapper of the CPU oracle (oracle/cgs_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.

`render_frame` chains the restated stages exactly as execute_frame_gpu does
(src/render/gpu_state.rs:613-1243) for a scene without cages.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_f = C.c_float
_i = C.c_int


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_set_threads": ([_i], None),
            "orc_max_threads": ([], _i),
            "orc_logf": ([_f], _f),
            "orc_expf": ([_f], _f),
            "orc_powf": ([_f, _f], _f),
            "orc_acosf": ([_f], _f),
            "orc_cosf": ([_f], _f),
            "orc_half_to_float": ([C.c_uint16], _f),
            "orc_project_v2": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _f,
                                _f, _f, _u32, _u32, _u32, _u32, _u32, _i, _vp, _vp], _u32),
            "orc_packed_key_bits": ([_u32, C.POINTER(_u32), C.POINTER(_u32)], None),
            "orc_count_tiles": ([_vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp],
                                None),
            "orc_exclusive_scan": ([_vp, _vp, _u64], _u32),
            "orc_emit_keys32": ([_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32,
                                 _u32, _vp, _vp], None),
            "orc_emit_keys64": ([_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp,
                                 _vp, _vp], None),
            "orc_sort_pairs32": ([_vp, _vp, _u64, _u32], None),
            "orc_sort64": ([_vp, _vp, _vp, _u64, _u32], None),
            "orc_tile_ranges": ([_vp, _u64, _u32, _u32, _vp], None),
            "orc_composite_v2": ([_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _vp, _i, _vp,
                                  _vp], None),
            "orc_mvc_far_share": ([_vp, _u32, _vp, _u32, _vp, _u64, _f, _f, _i, _vp], None),
            "orc_mvc_far_sums": ([_vp, _u32, _vp, _u32, _vp, _u64, _f, _f, _f, _i, _vp], None),
            "orc_mvc_scene": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _u32, _f, _f, _vp, _vp,
                               _vp, _vp], _u64),
            "orc_sh_tables": ([_vp], None),
            "orc_deformed_sh": ([_vp, _u32, _vp, _u32, _vp, _vp, _u32, _i, _vp], None),
            "orc_sh_band_matrix": ([_vp, _i, _vp], None),
            "orc_sh_rotate": ([_vp, _vp, _i], None),
            "orc_sh_rotate_fast": ([_vp, _vp, _i], None),
            "orc_sh_band_matrix_fast": ([_vp, _i, _vp], None),
            "orc_polar_rotation": ([_vp, _vp], None),
            "orc_sym_eigen": ([_vp, _vp, _vp], None),
            "orc_max_stretch": ([_vp], _f),
            "orc_split_major": ([_vp, _vp, _vp], None),
            "orc_mip_filter_3d": ([_vp, _vp, _f], None),
            "orc_eval_sh64": ([_vp, _vp, _u32, _vp], None),
            "orc_decode_activate": ([_vp, _u64, _u32, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                                    None),
            "orc_quantize": ([_vp, _vp, _u64, _u32, _vp, _vp, _vp], None),
            "orc_export_raw": ([_vp, _vp, _vp, _vp, _u64, _vp], None),
            "orc_f32_to_f16": ([_vp, _u64, _vp], None),
            "orc_atan2f": ([_f, _f], _f),
            "orc_quat_from_m3": ([_vp, _vp], None),
            "orc_encode_octahedral": ([_vp], _u32),
            "orc_decompose_covariance": ([_vp, _vp, _vp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_vp)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def packed_key_bits(total_tiles: int):
    tb, db = _u32(), _u32()
    lib().orc_packed_key_bits(total_tiles, C.byref(tb), C.byref(db))
    return tb.value, db.value


def base_depth_bits(near: float) -> int:
    return int(np.float32(max(near, 1e-6)).view(np.uint32)) & 0x7F800000


def project(scene, cam_arr, cam_pos, sh_degree=3, mip=0.0, min_opacity=1.0 / 255.0,
            min_radius=0.5, fp64=False, f16=None, bind_tet=None, tet_rot=None,
            split_meta=None, split_count=0, num_tets=0, max_gaussians=None, positions=None,
            covariances=None, opacities=None):
    """project_gaussians_v2_kernel restated; returns (attrs[n,10], extents[n], survivors).

    `f16` = (cov_halves[n,6] u16, sh_halves[n,stride] u16) selects the f16 variant.
    """
    n_rest = scene.count
    pos = positions if positions is not None else scene.positions
    cov = covariances if covariances is not None else scene.covariances_3d
    opa = opacities if opacities is not None else scene.opacities
    maxg = max_gaussians if max_gaussians is not None else n_rest
    attrs = np.zeros((maxg, 10), dtype=np.float64 if fp64 else np.float32)
    ext = np.zeros(maxg, dtype=np.uint32)
    cov_h = sh_h = None
    if f16 is not None:
        cov_h, sh_h = f16
    cam_pos = np.asarray(cam_pos, dtype=np.float32)
    has_def = 1 if bind_tet is not None else 0
    surv = lib().orc_project_v2(
        _p(np.ascontiguousarray(pos, dtype=np.float32)),
        None if cov_h is not None else _p(np.ascontiguousarray(cov, dtype=np.float32)),
        _p(cov_h), _p(np.ascontiguousarray(opa, dtype=np.float32)),
        None if sh_h is not None else _p(scene.sh_coeffs), _p(sh_h), _p(bind_tet), _p(tet_rot),
        _p(split_meta), split_count, _p(cam_arr), _p(cam_pos), mip, min_opacity, min_radius,
        sh_degree, has_def, num_tets, n_rest, maxg, 1 if fp64 else 0, _p(attrs), _p(ext))
    return attrs, ext, surv


def tile_grid(W, H, tile=(16, 16)):
    """TileGrid::new (tiles/mod.rs:39-49) for tiles of tile = (w, h) pixels."""
    tw, th = tile
    return (W + tw - 1) // tw, (H + th - 1) // th


def count_tiles(attrs, ext, W, H, active=None, tile=(16, 16)):
    n = ext.shape[0]
    active = n if active is None else active
    tx, ty = tile_grid(W, H, tile)
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_count_tiles(_p(np.ascontiguousarray(attrs, dtype=np.float32)), _p(ext), active, n,
                          W, H, tx, ty, tile[0], tile[1], _p(out))
    return out


def exclusive_scan(v):
    v = np.ascontiguousarray(v, dtype=np.uint32)
    out = np.zeros_like(v)
    tot = lib().orc_exclusive_scan(_p(v), _p(out), v.shape[0])
    return out, tot


def emit32(attrs, ext, offsets, total, W, H, near, active=None, tile=(16, 16)):
    n = ext.shape[0]
    active = n if active is None else active
    tx, ty = tile_grid(W, H, tile)
    _, db = packed_key_bits(tx * ty)
    keys = np.zeros(max(total, 1), dtype=np.uint32)
    ids = np.zeros(max(total, 1), dtype=np.uint32)
    lib().orc_emit_keys32(_p(np.ascontiguousarray(attrs, dtype=np.float32)), _p(ext), _p(offsets),
                          active, W, H, tx, ty, tile[0], tile[1], base_depth_bits(near), db,
                          _p(keys), _p(ids))
    return keys[:total], ids[:total], db


def emit64(attrs, ext, offsets, total, W, H, active=None, tile=(16, 16)):
    n = ext.shape[0]
    active = n if active is None else active
    tx, ty = tile_grid(W, H, tile)
    tiles = np.zeros(max(total, 1), dtype=np.uint32)
    depths = np.zeros(max(total, 1), dtype=np.uint32)
    ids = np.zeros(max(total, 1), dtype=np.uint32)
    lib().orc_emit_keys64(_p(np.ascontiguousarray(attrs, dtype=np.float32)), _p(ext), _p(offsets),
                          active, W, H, tx, ty, tile[0], tile[1], _p(tiles), _p(depths), _p(ids))
    return tiles[:total], depths[:total], ids[:total]


def sort_pairs32(keys, vals, bits=32):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    v = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    lib().orc_sort_pairs32(_p(k), _p(v), k.shape[0], bits)
    return k, v


def sort64(tiles, depths, ids, tile_bits):
    t = np.ascontiguousarray(tiles, dtype=np.uint32).copy()
    d = np.ascontiguousarray(depths, dtype=np.uint32).copy()
    i = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    lib().orc_sort64(_p(t), _p(d), _p(i), t.shape[0], tile_bits)
    return t, d, i


def tile_ranges(sorted_keys, shift, total_tiles):
    r = np.zeros(total_tiles * 2, dtype=np.uint32)
    s = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    lib().orc_tile_ranges(_p(s), s.shape[0], shift, total_tiles, _p(r))
    return r.reshape(-1, 2)


def composite(attrs, sorted_ids, ranges, W, H, bg, fp64=False, want_float=False, tile=(16, 16)):
    tx, _ = tile_grid(W, H, tile)
    pad = (W + 63) // 64 * 64
    out = np.zeros(pad * H, dtype=np.uint32)
    outf = np.zeros((H, W, 4), dtype=np.float32) if want_float else None
    a = np.ascontiguousarray(attrs, dtype=np.float64 if fp64 else np.float32)
    ids = np.ascontiguousarray(sorted_ids, dtype=np.uint32)
    if ids.size == 0:
        ids = np.zeros(1, dtype=np.uint32)
    lib().orc_composite_v2(_p(a), _p(ids), _p(np.ascontiguousarray(ranges, dtype=np.uint32)), W,
                           H, tx, tile[0], tile[1], pad, _p(np.asarray(bg, dtype=np.float32)),
                           1 if fp64 else 0, _p(out), _p(outf))
    rgba = out.reshape(H, pad)[:, :W].copy().view(np.uint8).reshape(H, W, 4)
    return (rgba, outf) if want_float else rgba


def render_frame(scene, cam, comp, sh_degree=None, mip=0.0, min_opacity=1.0 / 255.0,
                 min_radius=0.5, f16=None, tile=(16, 16)):
    """Stages 2-4 of execute_frame_gpu without cages, all intermediates returned;
    tile = (w, h) pixels per tile (the reference's 16 x 16 by default)."""
    from cindergs_amd.synthetic import camera_array  # plain data helper

    W, H = max(1, comp.screen_width), max(1, comp.screen_height)
    cam_arr = camera_array(cam)
    deg = comp.sh_degree if sh_degree is None else sh_degree
    attrs, ext, surv = project(scene, cam_arr, list(comp.camera_position), deg, mip, min_opacity,
                               min_radius, f16=f16)
    counts = count_tiles(attrs, ext, W, H, tile=tile)
    offsets, total = exclusive_scan(counts)
    tx, ty = tile_grid(W, H, tile)
    T = tx * ty
    tb, _ = packed_key_bits(T)
    out = {"attrs": attrs, "extents": ext, "survivors": surv, "counts": counts,
           "offsets": offsets, "total": total}
    if T <= 65535:
        keys, ids, db = emit32(attrs, ext, offsets, total, W, H, cam.near_plane, tile=tile)
        sk, sid = sort_pairs32(keys, ids, min(32, tb + db))
        ranges = tile_ranges(sk, db, T)
    else:
        tiles, depths, ids = emit64(attrs, ext, offsets, total, W, H, tile=tile)
        sk, _, sid = sort64(tiles, depths, ids, tb)
        ranges = tile_ranges(sk, 0, T)
    out.update(sorted_keys=sk, sorted_ids=sid, ranges=ranges)
    out["rgba"] = composite(attrs, sid, ranges, W, H, list(comp.background_color), tile=tile)
    return out


def deform(cage, scene, tau=0.0, mip=0.0, fp64=False):
    """Stage 1 restated (tet transforms + deform/split, children in parent order).

    Returns dict(J, R, stretch, pos[m,3], cov[m,6], opa[m], split_meta[s,2], splits)."""
    L = lib()
    if not hasattr(L, "_deform_sig"):
        L.orc_deform.argtypes = [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _f, _f, _i,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.orc_deform.restype = _u32
        L._deform_sig = True
    dt = np.float64 if fp64 else np.float32
    n = scene.count
    T = cage.elements.shape[0]
    J = np.zeros((T, 9), dtype=dt)
    R = np.zeros((T, 9), dtype=dt)
    st = np.zeros(T, dtype=dt)
    pos = np.zeros((2 * n, 3), dtype=dt)
    cov = np.zeros((2 * n, 6), dtype=dt)
    opa = np.zeros(2 * n, dtype=dt)
    meta = np.zeros((max(n, 1), 2), dtype=np.uint32)
    s = L.orc_deform(_p(np.ascontiguousarray(cage.vertices, dtype=np.float32)),
                     _p(np.ascontiguousarray(cage.elements, dtype=np.uint32)),
                     _p(np.ascontiguousarray(cage.inv_dm, dtype=np.float32)), T,
                     _p(np.ascontiguousarray(cage.bind_tet, dtype=np.uint32)),
                     _p(np.ascontiguousarray(cage.bind_weights, dtype=np.float32)),
                     _p(scene.positions), _p(scene.covariances_3d), _p(scene.opacities), n, tau,
                     mip, 1 if fp64 else 0, _p(J), _p(R), _p(st), _p(pos), _p(cov), _p(opa),
                     _p(meta))
    m = n + s
    return {"J": J, "R": R, "stretch": st, "pos": pos[:m], "cov": cov[:m], "opa": opa[:m],
            "split_meta": meta[:s], "splits": s}


def render_frame_deformed(scene, cage, cam, comp, tau=0.0, mip=0.0, min_opacity=1.0 / 255.0,
                          min_radius=0.5):
    """Stages 1-4 of execute_frame_gpu with a cage (Stage 2 mip = 0 with cages,
    gpu_state.rs:894; max_gaussians = 2N iff tau > 0, 706-713)."""
    from cindergs_amd.synthetic import camera_array

    d = deform(cage, scene, tau, mip)
    n = scene.count
    maxg = 2 * n if tau > 0 else n
    W, H = max(1, comp.screen_width), max(1, comp.screen_height)
    attrs, ext, surv = project(scene, camera_array(cam), list(comp.camera_position),
                               comp.sh_degree, 0.0, min_opacity, min_radius,
                               bind_tet=np.ascontiguousarray(cage.bind_tet, dtype=np.uint32),
                               tet_rot=np.ascontiguousarray(d["R"], dtype=np.float32),
                               split_meta=np.ascontiguousarray(d["split_meta"].reshape(-1)) if d["splits"] else np.zeros(2, np.uint32),
                               split_count=d["splits"], num_tets=cage.elements.shape[0],
                               max_gaussians=maxg, positions=np.ascontiguousarray(
                                   np.concatenate([d["pos"], np.zeros((maxg - d["pos"].shape[0], 3), np.float32)])),
                               covariances=np.ascontiguousarray(
                                   np.concatenate([d["cov"], np.zeros((maxg - d["cov"].shape[0], 6), np.float32)])),
                               opacities=np.ascontiguousarray(
                                   np.concatenate([d["opa"], np.zeros(maxg - d["opa"].shape[0], np.float32)])))
    active = n + d["splits"]
    counts = count_tiles(attrs, ext, W, H, active)
    offsets, total = exclusive_scan(counts)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    tb, _ = packed_key_bits(T)
    keys, ids, db = emit32(attrs, ext, offsets, total, W, H, cam.near_plane, active)
    sk, sid = sort_pairs32(keys, ids, min(32, tb + db))
    ranges = tile_ranges(sk, db, T)
    rgba = composite(attrs, sid, ranges, W, H, list(comp.background_color))
    return {"deform": d, "attrs": attrs, "extents": ext, "survivors": surv, "counts": counts,
            "offsets": offsets, "total": total, "sorted_keys": sk, "sorted_ids": sid,
            "ranges": ranges, "rgba": rgba}


def tet_inv_dm(verts, elems):
    """TetMesh::new's D_m^-1 per tet (mod.rs:202-227), column-major [T, 9] f32."""
    L = lib()
    if not hasattr(L, "_inv_sig"):
        L.orc_tet_inv_dm.argtypes = [_vp, _vp, _u32, _vp]
        L.orc_tet_inv_dm.restype = None
        L._inv_sig = True
    v = np.ascontiguousarray(verts, dtype=np.float32)
    e = np.ascontiguousarray(elems, dtype=np.uint32)
    out = np.zeros((e.shape[0], 9), np.float32)
    L.orc_tet_inv_dm(_p(v), _p(e), e.shape[0], _p(out))
    return out


def bind_tets(verts, elems, inv_dm, positions, bounds=None):
    """TetMesh::bind_gaussians(_with_bounds) (mod.rs:263-323), brute force in
    tet index order.  Returns (tet[n] u32 (0xFFFFFFFF unbound), weights[n,4])."""
    L = lib()
    if not hasattr(L, "_bind_sig"):
        L.orc_bind_tets.argtypes = [_vp, _vp, _vp, _u32, _vp, _u64, _vp, _vp, _vp]
        L.orc_bind_tets.restype = None
        L._bind_sig = True
    v = np.ascontiguousarray(verts, dtype=np.float32)
    e = np.ascontiguousarray(elems, dtype=np.uint32)
    m = np.ascontiguousarray(inv_dm, dtype=np.float32)
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    b = None if bounds is None else np.ascontiguousarray(
        np.concatenate([np.asarray(bounds[0], np.float32), np.asarray(bounds[1], np.float32)]))
    tet = np.zeros(p.shape[0], np.uint32)
    w = np.zeros((p.shape[0], 4), np.float32)
    L.orc_bind_tets(_p(v), _p(e), _p(m), e.shape[0], _p(p), p.shape[0], _p(b), _p(tet), _p(w))
    return tet, w


def mvc_bind(verts, tris, points, fp64=False, gradients=True):
    """PolyCageMesh::bind_gaussians (mvc.rs:498-541) restated: (weights [n,nv],
    gradients [n,nv,3] or None, inside [n] bool)."""
    L = lib()
    if not hasattr(L, "_mvc_sig"):
        L.orc_mvc_bind.argtypes = [_vp, _u32, _vp, _u32, _vp, _u64, _i, _vp, _vp, _vp]
        L.orc_mvc_bind.restype = None
        L.orc_mvc_winding.argtypes = [_vp, _vp, _u32, _vp, _u64, _i, _vp]
        L.orc_mvc_winding.restype = None
        L._mvc_sig = True
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    n, nv = p.shape[0], v.shape[0]
    dt = np.float64 if fp64 else np.float32
    w = np.zeros((n, nv), dt)
    g = np.zeros((n, nv, 3), dt) if gradients else None
    m = np.zeros(n, np.uint8)
    L.orc_mvc_bind(_p(v), nv, _p(t), t.shape[0], _p(p), n, 1 if fp64 else 0, _p(w), _p(g), _p(m))
    return w, g, m.astype(bool)


def mvc_far_share(verts, tris, points, r0, r1, fp64=True):
    """Far share of MVC's sums (triangle terms scaled by 1 - phi_t), times nv."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    p = np.ascontiguousarray(points, np.float32)
    out = np.zeros((p.shape[0], v.shape[0]), np.float64 if fp64 else np.float32)
    lib().orc_mvc_far_share(_p(v), v.shape[0], _p(t), t.shape[0], _p(p), p.shape[0], r0, r1,
                            1 if fp64 else 0, _p(out))
    return out


def mvc_far_sums(verts, tris, points, r0, r1, scale, fp64=True):
    """Far sums of MVC (triangle terms scaled by 1 - phi_t, not normalised),
    times scale."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    p = np.ascontiguousarray(points, np.float32)
    out = np.zeros((p.shape[0], v.shape[0]), np.float64 if fp64 else np.float32)
    lib().orc_mvc_far_sums(_p(v), v.shape[0], _p(t), t.shape[0], _p(p), p.shape[0], r0, r1,
                           float(scale), 1 if fp64 else 0, _p(out))
    return out


def far_sums_scale(verts):
    """The far sums' target scale of a cage: nv x bounding-box diagonal / (4 pi),
    in float32 as the library computes it."""
    v = np.asarray(verts, np.float32)
    d = (v.max(0) - v.min(0)).astype(np.float32)
    diag = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]).astype(np.float32)
    return float(np.float32(v.shape[0]) * max(diag, np.float32(1e-6)) * np.float32(0.0795774715))


def mvc_winding(verts, tris, points, fp64=False):
    """PolyCageMesh::winding_number (mvc.rs:326-347) per point (as float64)."""
    mvc_bind(verts, tris, np.zeros((0, 3), np.float32))  # signatures
    L = lib()
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.float64)
    L.orc_mvc_winding(_p(v), _p(t), t.shape[0], _p(p), p.shape[0], 1 if fp64 else 0, _p(out))
    return out


def mvc_weights(verts, tris, points, fp64=False):
    """PolyCageMesh::compute_mvc_weights (mvc.rs:366-471) without the inside test."""
    L = lib()
    if not hasattr(L, "_mvcw_sig"):
        L.orc_mvc_weights.argtypes = [_vp, _u32, _vp, _u32, _vp, _u64, _i, _vp]
        L.orc_mvc_weights.restype = None
        L._mvcw_sig = True
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.uint32)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    w = np.zeros((p.shape[0], v.shape[0]), np.float64 if fp64 else np.float32)
    L.orc_mvc_weights(_p(v), v.shape[0], _p(t), t.shape[0], _p(p), p.shape[0], 1 if fp64 else 0,
                      _p(w))
    return w


# ---- MVC scene deformation (deform_scene_cpu, mvc.rs:564-749) ----
# 3x3 matrices cross this boundary as numpy [3, 3] arrays; the library stores
# them column-major (float[9], column k = m[3k..3k+2]).

def _m3_in(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T).reshape(9)


def _m3_out(a):
    return a.reshape(3, 3).T.copy()


def mvc_scene(def_pos, def_F, mask, rest_pos, rest_cov, rest_opa, rest_sh, sh_degree=3,
              tau=0.0, mip=0.0):
    """PolyCageMesh::deform_scene_cpu after x' and F (def_F: [n, 9] column-major,
    as the library's Jacobians): (pos [m,3], cov [m,6], opa [m], sh [m,48])
    with m = n + splits, children appended in parent order."""
    n = int(np.asarray(rest_opa).shape[0])
    f = lambda a, k: np.ascontiguousarray(a, dtype=np.float32).reshape(n, k)
    dp, dF, rp, rc, rs = f(def_pos, 3), f(def_F, 9), f(rest_pos, 3), f(rest_cov, 6), f(rest_sh, 48)
    ro = np.ascontiguousarray(rest_opa, dtype=np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint32)
    op, oc = np.zeros((2 * n, 3), np.float32), np.zeros((2 * n, 6), np.float32)
    oo, osh = np.zeros(2 * n, np.float32), np.zeros((2 * n, 48), np.float32)
    total = lib().orc_mvc_scene(_p(dp), _p(dF), _p(m), _p(rp), _p(rc), _p(ro), _p(rs), n,
                                int(sh_degree), float(tau), float(mip), _p(op), _p(oc), _p(oo),
                                _p(osh))
    return op[:total], oc[:total], oo[:total], osh[:total]


def deformed_sh(rest_sh, rot, bind_tet, split_meta, degree):
    """SH of deform_scene_cpu's output (rest records, then split children in
    split_meta order); rot: [T, 9] column-major per-tet polar rotations."""
    rs = np.ascontiguousarray(rest_sh, dtype=np.float32).reshape(-1, 48)
    n = rs.shape[0]
    R = np.ascontiguousarray(rot, dtype=np.float32).reshape(-1, 9)
    bt = np.ascontiguousarray(bind_tet, dtype=np.uint32).reshape(n)
    meta = np.ascontiguousarray(split_meta, dtype=np.uint32).reshape(-1, 2)
    s = meta.shape[0]
    out = np.zeros((n + s, 48), np.float32)
    lib().orc_deformed_sh(_p(rs), n, _p(R), R.shape[0], _p(bt), _p(meta if s else None), s,
                          int(degree), _p(out))
    return out


def sh_tables():
    t = np.zeros(218, np.float64)
    lib().orc_sh_tables(_p(t))
    return t


def sh_band_matrix(R, l):
    """(2l+1) x (2l+1) rotation matrix of SH band l for the rotation R."""
    M = np.zeros((2 * l + 1) ** 2, np.float32)
    lib().orc_sh_band_matrix(_p(_m3_in(R)), int(l), _p(M))
    return M.reshape(2 * l + 1, 2 * l + 1)


def sh_band_matrix_fast(R, l):
    """The kernel's straight-line form of sh_band_matrix (cgs_sh_rotation_gen.h)."""
    M = np.zeros((2 * l + 1) ** 2, np.float32)
    lib().orc_sh_band_matrix_fast(_p(_m3_in(R)), int(l), _p(M))
    return M.reshape(2 * l + 1, 2 * l + 1)


def sh_rotate(sh, R, degree):
    out = np.ascontiguousarray(sh, dtype=np.float32).copy()
    lib().orc_sh_rotate(_p(out), _p(_m3_in(R)), int(degree))
    return out


def sh_rotate_fast(sh, R, degree):
    """The kernel's straight-line form of sh_rotate (cgs_sh_rotation_gen.h)."""
    out = np.ascontiguousarray(sh, dtype=np.float32).copy()
    lib().orc_sh_rotate_fast(_p(out), _p(_m3_in(R)), int(degree))
    return out


def polar_rotation(J):
    R = np.zeros(9, np.float32)
    lib().orc_polar_rotation(_p(_m3_in(J)), _p(R))
    return _m3_out(R)


def sym_eigen(cov):
    """(eigenvalues descending [3], basis [3,3] with the eigenvectors as columns)"""
    ev, B = np.zeros(3, np.float32), np.zeros(9, np.float32)
    lib().orc_sym_eigen(_p(np.ascontiguousarray(cov, dtype=np.float32)), _p(ev), _p(B))
    return ev, _m3_out(B)


def max_stretch(J):
    return float(lib().orc_max_stretch(_p(_m3_in(J))))


def split_major(cov):
    """(child covariance [6], offset [3]): children at -offset and +offset"""
    c, o = np.zeros(6, np.float32), np.zeros(3, np.float32)
    lib().orc_split_major(_p(np.ascontiguousarray(cov, dtype=np.float32)), _p(c), _p(o))
    return c, o


def mip_filter_3d(cov, alpha, var):
    c = np.ascontiguousarray(cov, dtype=np.float32).copy()
    a = np.array([alpha], np.float32)
    lib().orc_mip_filter_3d(_p(c), _p(a), float(var))
    return c, float(a[0])


def eval_sh64(coeffs, direction, degree):
    """Stage 2's SH colour in float64: clamp(0.5 + sum c Y(d), 0, 1) per channel."""
    out = np.zeros(3, np.float64)
    lib().orc_eval_sh64(_p(np.ascontiguousarray(coeffs, dtype=np.float64)),
                        _p(np.ascontiguousarray(direction, dtype=np.float64)), int(degree),
                        _p(out))
    return out


# ---- scene ingest (ply_loader.rs:540-598, soa.rs:113-126, 306-328, 454-500) ----

def decode_activate(payload, n, stride, little_endian, slot_offset, slot_type):
    """parse_vertex_payload + push_raw restated over a slot table: (raw [n,62],
    pos [n,3], cov [n,6], opa [n], sh [n,48])."""
    pl = np.frombuffer(bytes(payload), np.uint8)
    so = np.ascontiguousarray(slot_offset, dtype=np.uint32)
    st = np.ascontiguousarray(slot_type, dtype=np.uint8)
    raw = np.zeros((n, 62), np.float32)
    pos, cov = np.zeros((n, 3), np.float32), np.zeros((n, 6), np.float32)
    opa, sh = np.zeros(n, np.float32), np.zeros((n, 48), np.float32)
    lib().orc_decode_activate(_p(np.ascontiguousarray(pl)), n, stride, 1 if little_endian else 0,
                              _p(so), _p(st), _p(raw), _p(pos), _p(cov), _p(opa), _p(sh))
    return raw, pos, cov, opa, sh


def sh_words(degree):
    bases = {0: 1, 1: 4, 2: 9}.get(int(degree), 16)
    return (bases * 3 + 1) // 2


def quantize(cov, sh, degree):
    """from_soa_with_degree per splat: (cov_f16 [n,3] u32, rot [n] u32, sh_f16 [n,words] u32)"""
    c = np.ascontiguousarray(cov, dtype=np.float32).reshape(-1, 6)
    s = np.ascontiguousarray(sh, dtype=np.float32).reshape(-1, 48)
    n, w = c.shape[0], sh_words(degree)
    cw, r, sw = np.zeros((n, 3), np.uint32), np.zeros(n, np.uint32), np.zeros((n, w), np.uint32)
    lib().orc_quantize(_p(c), _p(s), n, int(degree), _p(cw), _p(r), _p(sw))
    return cw, r, sw


def export_raw(pos, cov, opa, sh):
    """to_raw_vertices: [n, 62] float records"""
    p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = p.shape[0]
    c = np.ascontiguousarray(cov, dtype=np.float32).reshape(n, 6)
    o = np.ascontiguousarray(opa, dtype=np.float32).reshape(n)
    s = np.ascontiguousarray(sh, dtype=np.float32).reshape(n, 48)
    raw = np.zeros((n, 62), np.float32)
    lib().orc_export_raw(_p(p), _p(c), _p(o), _p(s), n, _p(raw))
    return raw


def f32_to_f16_bits(x):
    a = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.zeros(a.shape[0], np.uint16)
    lib().orc_f32_to_f16(_p(a), a.shape[0], _p(out))
    return out


def atan2f(y, x):
    return float(lib().orc_atan2f(float(y), float(x)))


def quat_from_m3(m):
    """Quat::from_mat3 of a [3,3] matrix: (x, y, z, w)"""
    q = np.zeros(4, np.float32)
    lib().orc_quat_from_m3(_p(_m3_in(m)), _p(q))
    return q


def encode_octahedral(q_xyzw):
    return int(lib().orc_encode_octahedral(_p(np.ascontiguousarray(q_xyzw, dtype=np.float32))))


def decompose_covariance(cov):
    s, q = np.zeros(3, np.float32), np.zeros(4, np.float32)
    lib().orc_decompose_covariance(_p(np.ascontiguousarray(cov, dtype=np.float32)), _p(s), _p(q))
    return s, q


# ---------------------------------------------------------------------------
# O1: the reference's Rayon CPU render path (oracle/cgs_o1.cpp)
# ---------------------------------------------------------------------------

_O1_SIG = {
    "o1_fibonacci_cluster": ([_u32, _f, _vp, _vp, _vp, _vp, _vp], None),
    "o1_project": ([_vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _i], None),
    "o1_bin": ([_vp, _vp, _vp, _u64, _u32, _u32, _vp, _u64, _vp, _i], _u64),
    "o1_rasterize": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _u32, _vp, _vp, _i],
                     None),
    "o1_render": ([_vp, _vp, _vp, _vp, _u64, _vp, _u32, _u32, _vp, _vp, _u32, _vp, _vp, _vp, _i],
                  None),
    "o1_pixel_hash": ([_vp, _u64], _u64),
    "o1_evaluate_sh": ([_vp, _vp, _u32, _vp], None),
}


def _o1():
    L = lib()
    if not getattr(L, "_o1_bound", False):
        for name, (args, res) in _O1_SIG.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        L._o1_bound = True
    return L


def _threads(threads):
    return int(threads) if threads else (len(os.sched_getaffinity(0)) or 1)


def fibonacci_cluster(count: int, radius: float, center):
    """synthetic.rs:6-42 -> (positions[n,3], covariances[n,6], opacities[n], sh[n,48])."""
    pos = np.zeros((count, 3), np.float32)
    cov = np.zeros((count, 6), np.float32)
    opa = np.zeros(count, np.float32)
    sh = np.zeros((count, 48), np.float32)
    _o1().o1_fibonacci_cluster(count, radius, _p(np.asarray(center, np.float32)), _p(pos), _p(cov),
                               _p(opa), _p(sh))
    return pos, cov, opa, sh


def o1_camera(view16, fx, fy, cx, cy, w, h, near):
    """CameraUniforms::new (camera.rs:47-68) as the 24-float block."""
    a = np.zeros(24, np.float32)
    a[:16] = np.asarray(view16, np.float32).reshape(16)
    a[16:23] = (fx, fy, cx, cy, w, h, near)
    return a


def o1_project(pos, cov, cam_arr, threads=None):
    """project_scene_cpu (render/mod.rs:143-212) -> (points[n,2], depths, radii, conics[n,3])."""
    n = pos.shape[0]
    pts = np.zeros((n, 2), np.float32)
    d = np.zeros(n, np.float32)
    r = np.zeros(n, np.int32)
    c = np.zeros((n, 3), np.float32)
    _o1().o1_project(_p(np.ascontiguousarray(pos, np.float32)), _p(np.ascontiguousarray(cov, np.float32)), n,
                     _p(np.ascontiguousarray(cam_arr, np.float32)), _p(pts), _p(d), _p(r), _p(c),
                     _threads(threads))
    return pts, d, r, c


def o1_bin(pts, depths, radii, W, H, threads=None):
    """bin_and_sort_gaussians_cpu -> (entries[E,3] (tile, depth bits, id), ranges[T,2])."""
    n = radii.shape[0]
    T = ((W + 15) // 16) * ((H + 15) // 16)
    args = (_p(pts), _p(depths), _p(radii), n, W, H)
    total = _o1().o1_bin(*args, None, 0, None, _threads(threads))
    e = np.zeros((max(total, 1), 3), np.uint32)
    rng = np.zeros((T, 2), np.uint32)
    _o1().o1_bin(*args, _p(e), total, _p(rng), _threads(threads))
    return e[:total], rng


def o1_rasterize(pts, conics, opa, sh, pos, entries, ranges, W, H, bg, cam_pos, sh_degree,
                 threads=None):
    """rasterize_scene_cpu -> (rgba8[H,W,4], rgba_f32[H,W,4])."""
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgbaf = np.zeros((H, W, 4), np.float32)
    e = np.ascontiguousarray(entries, np.uint32)
    if e.size == 0:
        e = np.zeros((1, 3), np.uint32)
    _o1().o1_rasterize(_p(np.ascontiguousarray(pts, np.float32)), _p(np.ascontiguousarray(conics, np.float32)),
                       _p(np.ascontiguousarray(opa, np.float32)), _p(np.ascontiguousarray(sh, np.float32)),
                       _p(np.ascontiguousarray(pos, np.float32)), _p(e),
                       _p(np.ascontiguousarray(ranges, np.uint32)), W, H, _p(np.asarray(bg, np.float32)),
                       _p(np.asarray(cam_pos, np.float32)), sh_degree, _p(rgba8), _p(rgbaf),
                       _threads(threads))
    return rgba8, rgbaf


def o1_render(pos, cov, opa, sh, cam_arr, W, H, bg, cam_pos, sh_degree, threads=None):
    """render_scene_cpu (render/mod.rs:288-305).

    Returns (rgba8[H,W,4], {"project_ms", "bin_ms", "sort_ms", "raster_ms",
    "entries", "surviving"})."""
    n = pos.shape[0]
    rgba8 = np.zeros((H, W, 4), np.uint8)
    ms = np.zeros(4, np.float64)
    st = np.zeros(2, np.uint64)
    _o1().o1_render(_p(np.ascontiguousarray(pos, np.float32)), _p(np.ascontiguousarray(cov, np.float32)),
                    _p(np.ascontiguousarray(opa, np.float32)), _p(np.ascontiguousarray(sh, np.float32)), n,
                    _p(np.ascontiguousarray(cam_arr, np.float32)), W, H, _p(np.asarray(bg, np.float32)),
                    _p(np.asarray(cam_pos, np.float32)), sh_degree, _p(rgba8), _p(ms), _p(st),
                    _threads(threads))
    return rgba8, {"project_ms": float(ms[0]), "bin_ms": float(ms[1]), "sort_ms": float(ms[2]),
                   "raster_ms": float(ms[3]), "entries": int(st[0]), "surviving": int(st[1])}


def o1_render_scene(scene, cam, comp, threads=None):
    """O1 on a GaussianSceneSoa with the package's camera / compositor structs."""
    from cindergs_amd.synthetic import camera_array  # plain data helper

    return o1_render(scene.positions, scene.covariances_3d, scene.opacities, scene.sh_coeffs,
                     camera_array(cam), comp.screen_width, comp.screen_height,
                     list(comp.background_color), list(comp.camera_position), min(comp.sh_degree, 3),
                     threads)


def o1_evaluate_sh(coeffs48, direction, degree):
    """evaluate_sh (rasterizer/sh.rs:29-137) -> rgb[3]."""
    out = np.zeros(3, np.float32)
    _o1().o1_evaluate_sh(_p(np.ascontiguousarray(coeffs48, np.float32)),
                         _p(np.ascontiguousarray(direction, np.float32)), degree, _p(out))
    return out


def pixel_hash(rgba8) -> int:
    """RasterizedImage::pixel_hash (rasterizer/cpu_reference.rs:105-112)."""
    a = np.ascontiguousarray(rgba8, np.uint8)
    return int(_o1().o1_pixel_hash(_p(a), a.size))
