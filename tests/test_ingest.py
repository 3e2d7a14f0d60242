"""Scene ingest and quantised storage, CPU side.

* The library's PLY header parser (cgs_ply_parse_header: host code, no GPU)
  on the reference's ply_loader tests (src/scene/ply_loader.rs:747-1039):
  end-of-header detection (LF, CRLF, comments), vertex counts, ascii,
  missing properties, reordered / extra / big-endian properties, SH degree
  detection, compact SH scatter, unsupported SH layouts, list properties and
  extra elements; its slot table against the one the tests derive from the
  property list.
* The oracle's decode + activation (parse_vertex_payload + push_raw) against
  the tests' own struct unpacking, the reference's known answers
  (soa.rs:659-683, ply_loader.rs:805-1000) and a float64 closed form.
* The shared f32 -> f16 conversion against numpy's float16, atan2 against
  libm, Quat::from_mat3 against the rotation it must reproduce, the
  octahedral code by its decode round trip (soa.rs:615-644), and the export
  by reload (ply_loader.rs:1041-1080).
"""
import math

import numpy as np
import pytest

from ply_build import STANDARD, canonical_props, header_text, payload, random_rows, slot_table

F32 = np.float32


def _parse(data):
    from cindergs_amd.ply import parse_header

    return parse_header(data)


def _err(data):
    from cindergs_amd.ply import PlyError

    with pytest.raises(PlyError) as e:
        _parse(data)
    return e.value


# ---- header parser (ply_loader.rs:747-1039) ----

def test_end_of_header_offset_lf_crlf_and_comments():
    props = STANDARD
    lf = header_text("binary_little_endian 1.0", 0, props).encode()
    crlf = header_text("binary_little_endian 1.0", 0, props, eol="\r\n").encode()
    assert _parse(lf).payload_offset == len(lf)
    assert _parse(crlf).payload_offset == len(crlf)
    com = (b"ply\ncomment author: test\ncomment end_header in comment\n"
           b"format binary_little_endian 1.0\nelement vertex 0\nend_header\n")
    assert _parse(com).payload_offset == len(com)
    e = _err(b"ply\nformat binary_little_endian 1.0\nelement vertex 2\n")
    assert e.ply_kind == "MissingEndHeader"
    assert _err(b"plx\nend_header\n").ply_kind == "InvalidMagic"


def test_vertex_count_is_parsed():
    text = header_text("binary_little_endian 1.0", 14205, STANDARD).encode()
    assert _parse(text).vertex_count == 14205
    assert _err(b"ply\nformat binary_little_endian 1.0\nend_header\n").ply_kind == \
        "MissingVertexElement"


def test_ascii_is_rejected_with_a_clear_error():
    e = _err(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n")
    assert e.ply_kind == "UnsupportedFormat" and "ascii" in str(e)


def test_missing_required_property_is_rejected():
    text = header_text("binary_little_endian 1.0", 1, [("float", "x"), ("float", "y"),
                                                       ("float", "z")]).encode()
    e = _err(text)
    assert e.ply_kind == "MissingProperty" and "'f_dc_0'" in str(e)
    bad = [("int", "x")] + STANDARD[1:]
    e = _err(header_text("binary_little_endian 1.0", 1, bad).encode())
    assert e.ply_kind == "InvalidPropertyType" and "'x'" in str(e)


def test_sh_degree_detection():
    for rest, deg in ((0, 0), (9, 1), (24, 2), (45, 3)):
        props = STANDARD + [("float", f"f_rest_{i}") for i in range(rest)]
        assert _parse(header_text("binary_little_endian 1.0", 1, props).encode()).sh_degree == deg


def test_unsupported_sh_layout_is_rejected():
    props = STANDARD + [("float", "f_rest_0"), ("float", "f_rest_1")]
    e = _err(header_text("binary_little_endian 1.0", 1, props).encode())
    assert e.ply_kind == "UnsupportedShLayout" and e.detail[0] == 2


def test_list_properties_and_extra_elements_are_rejected():
    text = (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\n"
            b"property list uchar int vertex_indices\nend_header\n")
    e = _err(text)
    assert e.ply_kind == "UnsupportedListProperty" and "vertex_indices" in str(e)
    text = (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nelement face 2\n"
            b"property list uchar int vertex_indices\nproperty float x\nend_header\n")
    e = _err(text)
    assert e.ply_kind in ("UnsupportedListProperty", "UnsupportedElement")
    # an extra element with no records is accepted
    text = header_text("binary_little_endian 1.0", 0, STANDARD).encode().replace(
        b"end_header", b"element face 0\nend_header")
    assert _parse(text).vertex_count == 0


def test_slot_tables_of_the_reference_layouts():
    cases = {
        "canonical": canonical_props(),
        "reordered + int": [("float", "scale_0"), ("float", "scale_1"), ("float", "scale_2"),
                            ("float", "rot_0"), ("float", "rot_1"), ("float", "rot_2"),
                            ("float", "rot_3"), ("float", "opacity"), ("float", "f_dc_0"),
                            ("float", "f_dc_1"), ("float", "f_dc_2"), ("int", "cluster_id"),
                            ("float", "x"), ("float", "y"), ("float", "z")],
        "degree 1": STANDARD + [("float", f"f_rest_{i}") for i in range(9)],
        "degree 2 doubles": [("double", n) for _, n in STANDARD]
                            + [("float", f"f_rest_{i}") for i in range(24)]
                            + [("uchar", "red"), ("short", "tag")],
    }
    for name, props in cases.items():
        lay = _parse(header_text("binary_little_endian 1.0", 3, props).encode())
        so, st, stride = slot_table(props)
        assert lay.vertex_stride == stride, name
        np.testing.assert_array_equal(np.array(lay.slot_type), st, err_msg=name)
        np.testing.assert_array_equal(np.array(lay.slot_offset)[st > 0], so[st > 0], err_msg=name)
        assert lay.canonical == (1 if name == "canonical" else 0), name
    be = _parse(header_text("binary_big_endian 1.0", 1, canonical_props()).encode())
    assert be.format == 2 and be.canonical == 0


# ---- the oracle's decode + activation ----

def _decode(oracle_lib, props, rows, big_endian=False):
    so, st, stride = slot_table(props)
    return oracle_lib.decode_activate(payload(props, rows, big_endian), len(rows), stride,
                                      not big_endian, so, st)


def test_push_raw_identity_known_answer(oracle_lib):
    """soa.rs:659-683: opacity 0 -> 0.5, scale 0 -> unit covariance, identity rotation."""
    row = [1.0, 2.0, 3.0, 0, 0, 0, 0.1, 0.2, 0.3] + [0.0] * 45 + [0.0, 0.0, 0.0, 0.0,
                                                                 1.0, 0.0, 0.0, 0.0]
    raw, pos, cov, opa, sh = _decode(oracle_lib, canonical_props(), [row])
    assert pos[0].tolist() == [1.0, 2.0, 3.0] and opa[0] == 0.5
    assert cov[0].tolist() == [1.0, 0.0, 0.0, 1.0, 0.0, 1.0]
    np.testing.assert_array_equal(sh[0, :3], np.array([0.1, 0.2, 0.3], F32))
    # a zero quaternion falls back to the identity (covariance.rs:36-49)
    row[-4:] = [0.0, 0.0, 0.0, 0.0]
    assert _decode(oracle_lib, canonical_props(), [row])[2][0].tolist() == \
        [1.0, 0.0, 0.0, 1.0, 0.0, 1.0]


def test_reference_loader_cases(oracle_lib):
    # test_mmap_load_roundtrip (805-842)
    v = [10.0, -5.0, 3.5, 0, 0, 0, 0.5, 0.6, 0.7] + [0.0] * 45 + [2.0, -0.5, 0.0, 1.2,
                                                                   1.0, 0.0, 0.0, 0.0]
    raw, pos, cov, opa, sh = _decode(oracle_lib, canonical_props(), [v])
    assert pos[0].tolist() == [10.0, -5.0, 3.5]
    np.testing.assert_array_equal(sh[0, :3], np.array([0.5, 0.6, 0.7], F32))
    np.testing.assert_array_equal(raw[0], np.array(v, F32))
    # test_reordered_and_extra_properties_parse (868-910)
    props = [("float", "scale_0"), ("float", "scale_1"), ("float", "scale_2"),
             ("float", "rot_0"), ("float", "rot_1"), ("float", "rot_2"), ("float", "rot_3"),
             ("float", "opacity"), ("float", "f_dc_0"), ("float", "f_dc_1"),
             ("float", "f_dc_2"), ("int", "cluster_id"), ("float", "x"), ("float", "y"),
             ("float", "z")]
    row = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 3.0, 0.1, 0.2, 0.3, 7, 1.5, -2.0, 0.25]
    raw, pos, cov, opa, sh = _decode(oracle_lib, props, [row])
    assert pos[0].tolist() == [1.5, -2.0, 0.25]
    assert abs(opa[0] - 1.0 / (1.0 + math.exp(-3.0))) < 1e-6
    assert abs(sh[0, 0] - 0.1) < 1e-6
    # test_big_endian_payload_parses (912-933)
    row = [1.0, 2.0, 3.0, 0.1, 0.2, 0.3, 0.5, -1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    raw, pos, cov, opa, sh = _decode(oracle_lib, STANDARD, [row], big_endian=True)
    assert pos[0].tolist() == [1.0, 2.0, 3.0]
    np.testing.assert_array_equal(sh[0, :3], np.array([0.1, 0.2, 0.3], F32))
    # test_compact_sh_degree_one_is_scattered_per_channel (949-984)
    props = STANDARD + [("float", f"f_rest_{i}") for i in range(9)]
    raw, pos, cov, opa, sh = _decode(oracle_lib, props, [row + [float(i) for i in range(1, 10)]])
    s = sh[0]
    assert s[3:6].tolist() == [1.0, 2.0, 3.0] and s[18:21].tolist() == [4.0, 5.0, 6.0]
    assert s[33:36].tolist() == [7.0, 8.0, 9.0]
    assert not s[6:18].any() and not s[21:33].any() and not s[36:].any()


def test_decode_matches_struct_unpacking_on_every_scalar_type(oracle_lib):
    props = [("double", "x"), ("float", "y"), ("double", "z"), ("uchar", "nx"), ("char", "ny"),
             ("short", "nz")] + [(t, n) for t, n in STANDARD[3:]] + [("ushort", "junk"),
                                                                      ("uint", "id")]
    rows = random_rows(props, 40, 3)
    for be in (False, True):
        raw = _decode(oracle_lib, props, rows, big_endian=be)[0]
        names = [n for _, n in props]
        for r, row in enumerate(rows):
            for slot, name in ((0, "x"), (1, "y"), (2, "z"), (3, "nx"), (4, "ny"), (5, "nz"),
                               (54, "opacity"), (58, "rot_0")):
                want = np.float32(np.float64(row[names.index(name)]))
                assert raw[r, slot] == want, (be, r, name)


def _quat_to_mat(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_activation_matches_float64_closed_form(oracle_lib):
    props = canonical_props()
    rows = random_rows(props, 500, 11)
    raw, pos, cov, opa, sh = _decode(oracle_lib, props, rows)
    worst = 0.0
    for i in range(len(rows)):
        r = raw[i].astype(np.float64)
        s = np.maximum(np.exp(r[55:58]), 1e-4)
        q = r[[59, 60, 61, 58]]
        q = q / np.linalg.norm(q)
        R = _quat_to_mat(q)
        C = R @ np.diag(s * s) @ R.T
        ref = np.array([C[0, 0], C[0, 1], C[0, 2], C[1, 1], C[1, 2], C[2, 2]])
        worst = max(worst, float(np.abs(cov[i] - ref).max() / np.abs(ref).max()))
        assert abs(opa[i] - 1.0 / (1.0 + math.exp(-r[54]))) < 2e-7
    assert worst < 2e-6


# ---- f16, atan2, quaternions, octahedral code ----

def test_f16_conversion_matches_numpy_float16(oracle_lib):
    hi = np.arange(0, 1 << 16, dtype=np.uint32) << 16
    lows = np.array([0, 1, 0x0FFF, 0x1000, 0x1001, 0x2000, 0x3FFF, 0x8000, 0xFFFF], np.uint32)
    bits = (hi[:, None] | lows[None, :]).ravel()
    x = bits.view(np.float32)
    got = oracle_lib.f32_to_f16_bits(x)
    with np.errstate(over="ignore", invalid="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    fin = np.isfinite(x)
    np.testing.assert_array_equal(got[fin], ref[fin])
    nan = np.isnan(x)
    assert np.all((got[nan] & 0x7C00) == 0x7C00) and np.all((got[nan] & 0x03FF) != 0)
    assert np.all(got[np.isinf(x)] == ref[np.isinf(x)])


def test_atan2_matches_libm(oracle_lib):
    rng = np.random.default_rng(2)
    pts = [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0), (0.0, -0.0), (-0.0, -1.0),
           (1e-30, 1.0), (1.0, 1e-30), (3.0, -3.0), (0.41421356, 1.0)]
    pts += [tuple(v) for v in rng.normal(size=(2000, 2)) * np.exp(rng.uniform(-8, 8, (2000, 1)))]
    for y, x in pts:
        y32, x32 = float(np.float32(y)), float(np.float32(x))
        ref = math.atan2(y32, x32)
        got = oracle_lib.atan2f(y32, x32)
        assert abs(got - ref) <= 3e-7 * max(1.0, abs(ref)), (y, x, got, ref)


def _random_rotations(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q  # x, y, z, w


def test_quat_from_m3_reproduces_the_rotation(oracle_lib):
    for q in _random_rotations(300, 4):
        R = _quat_to_mat(q).astype(F32)
        qq = oracle_lib.quat_from_m3(R).astype(np.float64)
        assert abs(np.linalg.norm(qq) - 1.0) < 1e-6
        assert np.abs(_quat_to_mat(qq) - R).max() < 2e-6


def _decode_octahedral(packed):
    """decode_octahedral_rotation (soa.rs:404-423) in float64."""
    ox = ((packed & 0x7FF) / 2047.0) * 2.0 - 1.0
    oy = (((packed >> 11) & 0x7FF) / 2047.0) * 2.0 - 1.0
    oz = 1.0 - abs(ox) - abs(oy)
    if oz < 0.0:
        vx = (1.0 - abs(oy)) * (1.0 if ox >= 0 else -1.0)
        vy = (1.0 - abs(ox)) * (1.0 if oy >= 0 else -1.0)
    else:
        vx, vy = ox, oy
    axis = np.array([vx, vy, oz])
    n = np.linalg.norm(axis)
    axis = axis / n if n > 0 else np.array([0.0, 0.0, 1.0])
    angle = (((packed >> 22) & 0x3FF) / 1023.0) * math.pi
    return np.concatenate([axis * math.sin(angle / 2), [math.cos(angle / 2)]])


def _angle_between(a, b):
    d = min(1.0, abs(float(np.dot(a, b))))
    return 2.0 * math.acos(d)


def test_octahedral_rotation_round_trip(oracle_lib):
    """soa.rs:615-644 (Quat::from_euler(XYZ, 0.4, -0.7, 1.1), angle < 0.01),
    then 2000 random rotations."""
    def qx(a):
        return np.array([math.sin(a / 2), 0, 0, math.cos(a / 2)])

    def qy(a):
        return np.array([0, math.sin(a / 2), 0, math.cos(a / 2)])

    def qz(a):
        return np.array([0, 0, math.sin(a / 2), math.cos(a / 2)])

    def mul(a, b):
        ax, ay, az, aw = a
        bx, by, bz, bw = b
        return np.array([aw * bx + ax * bw + ay * bz - az * by,
                         aw * by - ax * bz + ay * bw + az * bx,
                         aw * bz + ax * by - ay * bx + az * bw,
                         aw * bw - ax * bx - ay * by - az * bz])

    r = mul(mul(qx(0.4), qy(-0.7)), qz(1.1))
    assert _angle_between(r, _decode_octahedral(oracle_lib.encode_octahedral(r))) < 0.01
    worst = 0.0
    for q in _random_rotations(2000, 9):
        worst = max(worst, _angle_between(q, _decode_octahedral(oracle_lib.encode_octahedral(q))))
    assert worst < 0.006
    ident = oracle_lib.encode_octahedral(np.array([0, 0, 0, 1.0], F32))
    assert (ident >> 22) == 0 and _angle_between([0, 0, 0, 1], _decode_octahedral(ident)) < 1e-6


def test_quantized_round_trip_and_compression(oracle_lib):
    """soa.rs:615-644: the quantised record reproduces covariance and SH to
    2e-3, compression > 1.8 at degree 3 and > 2.5 at degree 2."""
    from cindergs_amd.pipeline import QuantizedGaussianSceneSoa

    row = [0.5, -0.25, 1.5, 0, 0, 0, 0.4, -0.2, 0.6] + [0.0] * 45 + [1.2, -1.2, -1.8, -0.9,
                                                                     0.8, 0.3, -0.2, 0.4]
    row[9], row[9 + 15], row[9 + 30] = 0.15, -0.12, 0.08
    raw, pos, cov, opa, sh = _decode(oracle_lib, canonical_props(), [row])
    cw, rot, sw = oracle_lib.quantize(cov, sh, 3)
    unpacked = np.frombuffer(cw[0].astype("<u4").tobytes(), dtype="<f2").astype(np.float64)
    assert np.abs(unpacked - cov[0]).max() < 2e-3
    shv = np.frombuffer(sw[0].astype("<u4").tobytes(), dtype="<f2").astype(np.float64)
    rec = np.zeros(48)
    rec[:3] = shv[:3]
    for ch in range(3):
        rec[3 + 15 * ch:18 + 15 * ch] = shv[3 + 15 * ch:18 + 15 * ch]
    assert np.abs(rec - sh[0]).max() < 2e-3
    for deg, bound in ((3, 1.8), (2, 2.5)):
        cw, rot, sw = oracle_lib.quantize(cov, sh, deg)
        q = QuantizedGaussianSceneSoa(pos, cw, opa, sw, deg, rot)
        assert q.compression_ratio() > bound
    # the rotation code reproduces the principal axes of the covariance
    s, qd = oracle_lib.decompose_covariance(cov[0])
    R = _quat_to_mat(qd.astype(np.float64))
    C = R @ np.diag(s.astype(np.float64) ** 2) @ R.T
    ref = np.array([[cov[0][0], cov[0][1], cov[0][2]], [cov[0][1], cov[0][3], cov[0][4]],
                    [cov[0][2], cov[0][4], cov[0][5]]])
    assert np.abs(C - ref).max() < 1e-6


def test_export_then_reload_round_trip(oracle_lib):
    """ply_loader.rs:1041-1080 through the oracle: positions 1e-4,
    covariances 1e-3, opacities 1e-4, SH 1e-5."""
    pos = np.array([[0.25, -0.5, 1.25], [-1.0, 2.0, 0.5]], F32)
    cov = np.array([[0.04, 0.0, 0.0, 0.09, 0.0, 0.02], [0.01, 0.002, 0.0, 0.02, -0.001, 0.03]],
                   F32)
    opa = np.array([0.85, 0.35], F32)
    sh = np.zeros((2, 48), F32)
    sh[0, [0, 1, 2, 3, 18, 33]] = [0.5, -0.25, 0.75, 0.1, -0.15, 0.2]
    sh[1] = 0.12
    raw = oracle_lib.export_raw(pos, cov, opa, sh)
    data = (header_text("binary_little_endian 1.0", 2, canonical_props()).encode()
            + raw.tobytes())
    lay = _parse(data)
    assert lay.canonical == 1 and lay.vertex_count == 2
    so, st, stride = slot_table(canonical_props())
    _, p2, c2, o2, s2 = oracle_lib.decode_activate(data[lay.payload_offset:], 2, stride, True,
                                                   so, st)
    assert np.abs(p2 - pos).max() < 1e-4
    assert np.abs(c2 - cov).max() < 1e-3
    assert np.abs(o2 - opa).max() < 1e-4
    assert abs(s2[0, 0] - 0.5) < 1e-5 and abs(s2[0, 33] - 0.2) < 1e-5
