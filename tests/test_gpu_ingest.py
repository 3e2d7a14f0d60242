"""Scene ingest and quantised storage on the GPU against the oracle.

cgs_ply_load (header parsed on the host, payload decoded and activated by
k_activate_raw) is compared bit for bit with the oracle's parse_vertex_payload
+ push_raw on the same bytes, for the canonical 3DGS record and for every
other layout the reference accepts: reordered and extra properties, integer
properties, doubles, big-endian payloads, compact SH of degree 0-2, CRLF
headers, payloads at any byte offset.  cgs_quantize_scene and cgs_export_raw
are compared bit for bit with the oracle's from_soa_with_degree and
to_raw_vertices; the reference's own loader cases run through the GPU path.
"""
import zlib

import numpy as np
import pytest

from ply_build import (STANDARD, canonical_props, header_text, payload, ply_bytes, random_rows,
                       slot_table)

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _load(gpu, data):
    from cindergs_amd.ply import PlyLoader

    n, out = PlyLoader.load_to_device(gpu, data)
    got = [out[k].numpy()[:n] for k in ("pos", "cov", "opa", "sh")]
    for b in out.values():
        b.free()
    return n, got


def _oracle(oracle_lib, data, props, big_endian=False):
    from cindergs_amd.ply import parse_header

    lay = parse_header(data)
    so, st, stride = slot_table(props)
    n = int(lay.vertex_count)
    return oracle_lib.decode_activate(data[lay.payload_offset:lay.payload_offset + n * stride],
                                      n, stride, not big_endian, so, st)[1:]


LAYOUTS = {
    "canonical": (canonical_props(), False),
    "reordered + int": ([("float", "scale_0"), ("float", "scale_1"), ("float", "scale_2"),
                         ("float", "rot_0"), ("float", "rot_1"), ("float", "rot_2"),
                         ("float", "rot_3"), ("float", "opacity"), ("float", "f_dc_0"),
                         ("float", "f_dc_1"), ("float", "f_dc_2"), ("int", "cluster_id"),
                         ("float", "x"), ("float", "y"), ("float", "z")], False),
    "big-endian canonical": (canonical_props(), True),
    "degree 1 + bytes": (STANDARD + [("float", f"f_rest_{i}") for i in range(9)]
                         + [("uchar", "red"), ("char", "tag")], False),
    "degree 2 doubles": ([("double", n) for _, n in STANDARD]
                         + [("float", f"f_rest_{i}") for i in range(24)]
                         + [("short", "s"), ("ushort", "u")], False),
    "degree 0 big-endian mixed": ([("double", "x"), ("float", "y"), ("double", "z"),
                                   ("uint", "id"), ("float", "nx")] + STANDARD[3:], True),
}


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_load_matches_oracle_bit_exact(gpu, oracle_lib, name):
    props, be = LAYOUTS[name]
    n = 1031 if name == "canonical" else 517
    rows = random_rows(props, n, zlib.crc32(name.encode()) % 1000)
    data = ply_bytes(props, rows, big_endian=be)
    print(f"[case] ply load {name} n={n} bytes={len(data)}", flush=True)
    count, got = _load(gpu, data)
    ref = _oracle(oracle_lib, data, props, be)
    assert count == n
    for k, a, b in zip(("pos", "cov", "opa", "sh"), got, ref):
        np.testing.assert_array_equal(_bits(a), _bits(b), err_msg=f"{name} {k}")


def test_payload_at_any_offset_and_crlf(gpu, oracle_lib):
    props = canonical_props()
    rows = random_rows(props, 300, 5)
    for pad in range(4):
        text = header_text("binary_little_endian 1.0", len(rows), props, eol="\r\n")
        text = text.replace("format", "comment " + "x" * pad + "\r\nformat", 1)
        data = text.encode() + payload(props, rows)
        print(f"[case] ply load crlf pad={pad}", flush=True)
        count, got = _load(gpu, data)
        ref = _oracle(oracle_lib, data, props)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(_bits(a), _bits(b))


def test_reference_loader_cases_on_gpu(gpu):
    import math

    print("[case] ply reference cases", flush=True)
    v = [10.0, -5.0, 3.5, 0, 0, 0, 0.5, 0.6, 0.7] + [0.0] * 45 + [2.0, -0.5, 0.0, 1.2,
                                                                   1.0, 0.0, 0.0, 0.0]
    n, (pos, cov, opa, sh) = _load(gpu, ply_bytes(canonical_props(), [v]))
    assert n == 1 and pos[0].tolist() == [10.0, -5.0, 3.5]
    assert sh[0, :3].tolist() == np.float32([0.5, 0.6, 0.7]).tolist()
    props = LAYOUTS["reordered + int"][0]
    row = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 3.0, 0.1, 0.2, 0.3, 7, 1.5, -2.0, 0.25]
    n, (pos, cov, opa, sh) = _load(gpu, ply_bytes(props, [row]))
    assert pos[0].tolist() == [1.5, -2.0, 0.25]
    assert abs(opa[0] - 1.0 / (1.0 + math.exp(-3.0))) < 1e-6
    row = [1.0, 2.0, 3.0, 0.1, 0.2, 0.3, 0.5, -1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    n, (pos, cov, opa, sh) = _load(gpu, ply_bytes(STANDARD, [row], big_endian=True))
    assert pos[0].tolist() == [1.0, 2.0, 3.0]
    props = STANDARD + [("float", f"f_rest_{i}") for i in range(9)]
    n, (pos, cov, opa, sh) = _load(gpu, ply_bytes(props, [row + [float(i) for i in range(1, 10)]]))
    s = sh[0]
    assert s[3:6].tolist() == [1.0, 2.0, 3.0] and s[18:21].tolist() == [4.0, 5.0, 6.0]
    assert s[33:36].tolist() == [7.0, 8.0, 9.0] and not s[36:].any()


def test_truncated_empty_and_errors(gpu):
    from cindergs_amd.ply import PlyError, PlyLoader

    print("[case] ply truncated / empty", flush=True)
    text = header_text("binary_little_endian 1.0", 2, STANDARD).encode()
    with pytest.raises(PlyError) as e:
        PlyLoader.load_to_device(gpu, text + bytes(8))
    assert e.value.ply_kind == "TruncatedPayload" and e.value.detail == (2 * 14 * 4, 8)
    # a vertex count whose payload size wraps 64 bits is truncated, not a
    # wrapped small payload
    huge = (1 << 64) // (14 * 4) + 1
    text = header_text("binary_little_endian 1.0", huge, STANDARD).encode()
    with pytest.raises(PlyError) as e:
        PlyLoader.load_to_device(gpu, text + bytes(64))
    assert e.value.ply_kind == "TruncatedPayload" and e.value.detail[1] == 64
    n, out = PlyLoader.load_to_device(gpu, header_text("binary_little_endian 1.0", 0,
                                                       STANDARD).encode())
    assert n == 0
    for b in out.values():
        b.free()
    with pytest.raises(PlyError):
        PlyLoader.load_to_device(gpu, b"ply\nformat ascii 1.0\nelement vertex 1\nend_header\n")


def _scene(n, seed):
    from cindergs_amd import synthetic as syn

    return syn.scene_gaussian_ball(n, k=seed)


@pytest.mark.parametrize("degree", [3, 2, 1, 0])
def test_quantize_matches_oracle_bit_exact(gpu, oracle_lib, degree):
    from cindergs_amd.ply import quantize_scene

    scene = _scene(20000, 7)
    print(f"[case] quantize degree {degree}", flush=True)
    q = quantize_scene(gpu, scene, degree)
    cw, rot, sw = oracle_lib.quantize(scene.covariances_3d, scene.sh_coeffs, degree)
    np.testing.assert_array_equal(q.covariances_f16_packed, cw)
    np.testing.assert_array_equal(q.octahedral_rotations, rot)
    np.testing.assert_array_equal(q.sh_coeffs_f16_packed, sw)


def test_export_matches_oracle_and_reloads(gpu, oracle_lib):
    from cindergs_amd.ply import PlyLoader

    scene = _scene(20000, 8)
    print("[case] export + reload", flush=True)
    data = PlyLoader.write_binary_ply_bytes(scene, gpu)
    from cindergs_amd.ply import parse_header

    lay = parse_header(data)
    assert lay.canonical == 1 and lay.vertex_count == scene.count
    raw = np.frombuffer(data[lay.payload_offset:], dtype=np.float32).reshape(-1, 62)
    ref = oracle_lib.export_raw(scene.positions, scene.covariances_3d, scene.opacities,
                                scene.sh_coeffs)
    np.testing.assert_array_equal(_bits(raw), _bits(ref))
    back = PlyLoader.load_from_bytes(data, gpu)
    assert np.abs(back.positions - scene.positions).max() == 0.0
    scale = np.abs(scene.covariances_3d).max(axis=1, keepdims=True)
    assert np.all(np.abs(back.covariances_3d - scene.covariances_3d) <= 1e-4 * scale)
    assert np.abs(back.opacities - scene.opacities).max() < 1e-5
    np.testing.assert_array_equal(back.sh_coeffs, scene.sh_coeffs)
