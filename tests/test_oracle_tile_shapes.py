"""The oracle's Stage 3 and Stage 4 with the tile size as a parameter
(count_tiles / emit32 / emit64 / composite with tile = (w, h)), pinned at every
size the library offers by brute force on small screens:

* the tiles a splat is emitted for are exactly those of its extent box
  (compute_gaussian_tile_bounds_xy, tiles/mod.rs:97-130, with w, h for 16)
  whose pixel-centre rectangle the splat's conic reaches within its cutoff
  (ellipse_intersects_tile_gpu, binning.rs:35-83): the minimum of the
  quadratic over the rectangle, found here in float64 over all four edges (the
  reference checks the two edges facing the centre, which hold the minimum),
  against clamp(2 ln(255 alpha), 0, 8) + 1e-4; pairs within 1e-5 of the
  threshold are left out;
* no tile holding a pixel centre the compositor blends is missed (the
  reference's own test, tiles/mod.rs:302-401, at each size);
* the frame equals a per-pixel front-to-back blend of every splat in the order
  of the size's sort keys (tile << depth_bits | depth, depth_bits = 32 - tile
  bits, binning.rs:12-20, then splat id), with the background's alpha where
  the pixel's tile holds no entry (compositor.rs:721-883);
* 16 x 16 gives the same keys as the default arguments, and a frame of more
  than 65535 tiles takes the 64-bit keys (tile, depth bits) with the same
  frame as the brute-force blend in (tile, depth, id) order.
"""
import math

import numpy as np
import pytest

SHAPES = [(8, 8), (16, 16), (16, 32), (32, 16)]
F32 = np.float32


def _splats(n, W, H, seed, depth_levels=None):
    """Random projected splats: attrs10 rows (mu_x, mu_y, a, b, c, alpha, r,
    g, b, depth) and packed extents (ceil(sigma sqrt(Sigma_xx)) | ... << 16)."""
    rng = np.random.default_rng(seed)
    rows, ext = [], []
    while len(rows) < n:
        mu_x = F32(rng.uniform(-20, W + 20))
        mu_y = F32(rng.uniform(-20, H + 20))
        ang = rng.uniform(0, math.pi)
        l1, l2 = rng.uniform(0.5, 400.0), rng.uniform(0.5, 20.0)
        s, c = math.sin(ang), math.cos(ang)
        sxx = F32(c * c * l1 + s * s * l2 + 0.3)
        sxy = F32(s * c * (l1 - l2))
        syy = F32(s * s * l1 + c * c * l2 + 0.3)
        alpha = F32(rng.uniform(0.01, 0.99))
        det = sxx * syy - sxy * sxy
        if det <= F32(1e-6):
            continue
        sig = math.sqrt(min(max(2.0 * math.log(255.0 * float(alpha)), 0.0), 8.0))
        ex = int(min(max(math.ceil(sig * math.sqrt(float(sxx))), 1), 1024))
        ey = int(min(max(math.ceil(sig * math.sqrt(float(syy))), 1), 1024))
        if mu_x + ex < 0 or mu_x - ex >= W or mu_y + ey < 0 or mu_y - ey >= H:
            continue  # projection culls boxes entirely off screen
        inv = F32(1.0) / det
        if depth_levels is None:
            depth = F32(rng.uniform(0.5, 6.0))
        else:  # near-equal depths: the sort's depth quantization decides their order
            depth = F32(depth_levels[rng.integers(len(depth_levels))] * (1 + rng.uniform(0, 3e-5)))
        rows.append([mu_x, mu_y, syy * inv, -sxy * inv, sxx * inv, alpha,
                     rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1), depth])
        ext.append(ex | (ey << 16))
    return (np.ascontiguousarray(np.asarray(rows, np.float32)),
            np.ascontiguousarray(np.asarray(ext, np.uint32)))


def _rect_min(a, b, c, x0, x1, y0, y1):
    """min of a x^2 + 2 b x y + c y^2 (positive definite) over [x0,x1] x [y0,y1], float64."""
    if x0 <= 0.0 <= x1 and y0 <= 0.0 <= y1:
        return 0.0
    best = math.inf
    for xe in (x0, x1):
        y = min(max(-b * xe / c, y0), y1)
        best = min(best, a * xe * xe + 2 * b * xe * y + c * y * y)
    for ye in (y0, y1):
        x = min(max(-b * ye / a, x0), x1)
        best = min(best, a * x * x + 2 * b * x * ye + c * ye * ye)
    return best


def _brute_tiles(attrs, ext, W, H, tile):
    """{splat: set of tiles} decided by the float64 minimum, and the set of
    (splat, tile) pairs too close to the threshold to decide."""
    tw, th = tile
    tx_n, ty_n = (W + tw - 1) // tw, (H + th - 1) // th
    sure, unsure = {}, set()
    for i, r in enumerate(attrs.astype(np.float64)):
        mu_x, mu_y, a, b, c, alpha = r[:6]
        exx, exy = float(ext[i] & 0xFFFF), float(ext[i] >> 16)
        fx0 = int(np.floor(F32(mu_x - exx) / F32(tw)))
        fx1 = int(np.floor(F32(mu_x + exx) / F32(tw)))
        fy0 = int(np.floor(F32(mu_y - exy) / F32(th)))
        fy1 = int(np.floor(F32(mu_y + exy) / F32(th)))
        bx0, bx1 = min(max(fx0, 0), tx_n - 1), min(max(fx1, 0), tx_n - 1)
        by0, by1 = min(max(fy0, 0), ty_n - 1), min(max(fy1, 0), ty_n - 1)
        thr = min(max(2.0 * math.log(255.0 * alpha), 0.0), 8.0) + 1e-4
        hit = set()
        if alpha >= 1.0 / 255.0:
            for ty in range(by0, by1 + 1):
                for tx in range(bx0, bx1 + 1):
                    xs, ys = tx * tw, ty * th
                    xe, ye = min(xs + tw - 1, W - 1), min(ys + th - 1, H - 1)
                    m = _rect_min(a, b, c, xs + 0.5 - mu_x, xe + 0.5 - mu_x, ys + 0.5 - mu_y,
                                  ye + 0.5 - mu_y)
                    if abs(m - thr) <= 1e-5 * (1.0 + thr):
                        unsure.add((i, ty * tx_n + tx))
                    elif m <= thr:
                        hit.add(ty * tx_n + tx)
        sure[i] = hit
    return sure, unsure


def _emitted(o, attrs, ext, W, H, tile):
    counts = o.count_tiles(attrs, ext, W, H, tile=tile)
    offs, total = o.exclusive_scan(counts)
    keys, ids, db = o.emit32(attrs, ext, offs, total, W, H, 0.2, tile=tile)
    got = {}
    for k, i in zip(keys, ids):
        got.setdefault(int(i), set()).add(int(k) >> db)
    return counts, got, keys, ids, db


@pytest.mark.parametrize("tile", SHAPES)
def test_emitted_tiles_are_exactly_the_reached_tiles(oracle_lib, tile):
    o = oracle_lib
    W, H = 150, 118  # neither a multiple of 32: partial tiles on both edges
    attrs, ext = _splats(400, W, H, seed=11 + tile[0] + 3 * tile[1])
    counts, got, _, _, _ = _emitted(o, attrs, ext, W, H, tile)
    sure, unsure = _brute_tiles(attrs, ext, W, H, tile)
    assert len(unsure) <= 2
    for i in range(attrs.shape[0]):
        g = got.get(i, set())
        assert counts[i] == len(g)
        amb = {t for (j, t) in unsure if j == i}
        assert g - amb == sure[i] - amb, (tile, i, sorted(g ^ sure[i]))


@pytest.mark.parametrize("tile", SHAPES)
def test_no_blended_pixel_is_missed(oracle_lib, tile):
    """tiles/mod.rs:302-401 at each tile size: every tile holding a pixel
    centre the compositor would blend (q >= -4, alpha >= 1/255) is emitted."""
    o = oracle_lib
    tw, th = tile
    W, H = 160, 128
    tx_n = W // tw
    attrs, ext = _splats(300, W, H, seed=5 + tw)
    _, got, _, _, _ = _emitted(o, attrs, ext, W, H, tile)
    px = (np.arange(W, dtype=np.float32) + F32(0.5))[None, :]
    py = (np.arange(H, dtype=np.float32) + F32(0.5))[:, None]
    for i in range(attrs.shape[0]):
        mx, my, a, b, c, al = (F32(v) for v in attrs[i, :6])
        dx, dy = px - mx, py - my
        q = F32(-0.5) * (a * dx * dx + F32(2.0) * b * dx * dy + c * dy * dy)
        hit = (q >= F32(-4.0)) & (np.minimum(F32(0.99), al * np.exp(q)) >= F32(1.0 / 255.0))
        hit_t = hit.reshape(H // th, th, W // tw, tw).any(axis=(1, 3))
        for ty, tx in zip(*np.nonzero(hit_t)):
            assert ty * tx_n + tx in got.get(i, set()), (tile, i, tx, ty)


def _blend_all(attrs, order, W, H, bg, tile_of_pixel_empty):
    """Front-to-back blend of the splats in `order` at every pixel (float32,
    the compositor's per-pixel steps), then the RGBA8 rounding."""
    px = (np.arange(W, dtype=np.float32) + F32(0.5))[None, :]
    py = (np.arange(H, dtype=np.float32) + F32(0.5))[:, None]
    C = np.zeros((3, H, W), np.float32)
    T = np.ones((H, W), np.float32)
    alive = np.ones((H, W), bool)
    for i in order:
        mx, my, a, b, c, op = (F32(v) for v in attrs[i, :6])
        if not op > F32(0.003921569):
            continue
        dx, dy = px - mx, py - my
        q = F32(-0.5) * (a * dx * dx + F32(2.0) * b * dx * dy + c * dy * dy)
        q = np.minimum(q, F32(0.0))
        al = np.minimum(F32(0.99), op * np.exp(q).astype(np.float32))
        ok = alive & (q >= F32(-4.0)) & (al >= F32(0.003921569))
        w = np.where(ok, al * T, F32(0.0)).astype(np.float32)
        for ch in range(3):
            C[ch] += w * F32(attrs[i, 6 + ch])
        Tn = np.where(ok, T * (F32(1.0) - al), T).astype(np.float32)
        sat = ok & (Tn < F32(0.001))
        T = np.where(sat, F32(0.0), Tn)
        alive &= ~sat
    out = np.zeros((H, W, 4), np.float32)
    for ch in range(3):
        out[..., ch] = np.clip(C[ch] + T * F32(bg[ch]), 0, 1)
    out[..., 3] = np.where(tile_of_pixel_empty, F32(bg[3]), np.clip(F32(1.0) - T, 0, 1))
    return np.clip(np.round(out * 255), 0, 255).astype(np.int16)


@pytest.mark.parametrize("tile", SHAPES)
def test_frame_is_the_blend_in_the_sizes_key_order(oracle_lib, tile):
    o = oracle_lib
    tw, th = tile
    W, H = 150, 118
    bg = (0.1, 0.2, 0.3, 1.0)
    # few depth levels, each spread over ~3e-5 relative (~250 ulps): the key's
    # depth bits (which depend on the tile count) decide the order inside a level
    attrs, ext = _splats(250, W, H, seed=77 + tw * th, depth_levels=[0.7, 1.3, 2.9])
    counts, got, keys, ids, db = _emitted(o, attrs, ext, W, H, tile)
    tx_n, ty_n = o.tile_grid(W, H, tile)
    tb, db2 = o.packed_key_bits(tx_n * ty_n)
    assert db2 == db == min(32 - tb, 28)
    sk, sid = o.sort_pairs32(keys, ids, min(32, tb + db))
    ranges = o.tile_ranges(sk, db, tx_n * ty_n)
    rgba = o.composite(attrs, sid, ranges, W, H, bg, tile=tile).astype(np.int16)
    # the depth part of each splat's key, then the blend order (depth part, id)
    base = o.base_depth_bits(0.2)
    raw = np.maximum(attrs[:, 9], F32(1e-6)).view(np.uint32).astype(np.int64)
    dq = np.minimum(np.maximum(raw - base, 0), 0x0FFFFFFF) >> (28 - db)
    order = np.lexsort((np.arange(attrs.shape[0]), dq))
    # RGBA alpha of a pixel whose tile has no entry is the background's
    occupied = np.zeros(tx_n * ty_n, bool)
    occupied[[t for s in got.values() for t in s]] = True
    ty_of = np.arange(H) // th
    tx_of = np.arange(W) // tw
    empty = ~occupied[ty_of[:, None] * tx_n + tx_of[None, :]]
    ref = _blend_all(attrs, order, W, H, bg, empty)
    d = np.abs(rgba - ref)
    assert d.max() <= 1, (tile, int(d.max()))
    if tile == (8, 8):
        # the data exercise the size's own depth quantization: the 16 x 16
        # grid's depth bits would order these splats differently
        _, db16 = o.packed_key_bits(((W + 15) // 16) * ((H + 15) // 16))
        dq16 = np.minimum(np.maximum(raw - base, 0), 0x0FFFFFFF) >> (28 - db16)
        order16 = np.lexsort((np.arange(attrs.shape[0]), dq16))
        assert not np.array_equal(order, order16)


def test_16x16_is_the_default(oracle_lib):
    o = oracle_lib
    W, H = 150, 118
    attrs, ext = _splats(200, W, H, seed=3)
    c0 = o.count_tiles(attrs, ext, W, H)
    c1 = o.count_tiles(attrs, ext, W, H, tile=(16, 16))
    np.testing.assert_array_equal(c0, c1)
    offs, total = o.exclusive_scan(c0)
    k0, i0, _ = o.emit32(attrs, ext, offs, total, W, H, 0.2)
    k1, i1, _ = o.emit32(attrs, ext, offs, total, W, H, 0.2, tile=(16, 16))
    np.testing.assert_array_equal(k0, k1)
    np.testing.assert_array_equal(i0, i1)


def test_64bit_keys_at_8x8(oracle_lib):
    """2080 x 2080 at 8 x 8 is 67600 tiles (> 65535): (tile, depth bits, id)
    keys; the frame is the blend in (depth, id) order, as no splat is cut by
    the depth quantization."""
    o = oracle_lib
    W = H = 2080
    tile = (8, 8)
    attrs, ext = _splats(60, 200, 200, seed=9)
    attrs[:, 0] += F32(900.0)
    attrs[:, 1] += F32(900.0)
    tx_n, ty_n = o.tile_grid(W, H, tile)
    assert tx_n * ty_n > 65535
    counts = o.count_tiles(attrs, ext, W, H, tile=tile)
    offs, total = o.exclusive_scan(counts)
    tiles, depths, ids = o.emit64(attrs, ext, offs, total, W, H, tile=tile)
    tb, _ = o.packed_key_bits(tx_n * ty_n)
    st, sd, sid = o.sort64(tiles, depths, ids, tb)
    order = np.lexsort((sid, sd, st))
    np.testing.assert_array_equal(order, np.arange(order.shape[0]))
    ranges = o.tile_ranges(st, 0, tx_n * ty_n)
    bg = (0.0, 0.0, 0.0, 0.0)
    rgba = o.composite(attrs, sid, ranges, W, H, bg, tile=tile).astype(np.int16)
    # the splats reach at most 60 px past [880, 1120]: the window [720, 1280)
    # holds them; the rest is background (alpha 0, as is an unreached pixel)
    lo, hi = 720, 1280
    sub = attrs.copy()
    sub[:, 0] -= F32(lo)  # exact: multiples of 8 keep the pixel centres' bits
    sub[:, 1] -= F32(lo)
    blend_order = np.lexsort((np.arange(attrs.shape[0]), attrs[:, 9]))
    ref = _blend_all(sub, blend_order, hi - lo, hi - lo, bg, np.zeros((hi - lo, hi - lo), bool))
    assert np.abs(rgba[lo:hi, lo:hi] - ref).max() <= 1
    rgba[lo:hi, lo:hi] = 0
    assert not rgba.any()
