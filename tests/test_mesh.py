"""Mesh ingestion (pppm_mesh_load and the C++ Trimesh): CPU only.

* On the reference's own assets (when the reference tree is present) the
  loaded and placed triangles equal those of the compiled reference's
  Trimesh (load, fitToBBox, scale, putOnPlane, translate as its application
  calls them) bit for bit.
* Files written here in every layout the loader accepts (binary little and
  big endian, ASCII, extra vertex properties, double coordinates, polygons,
  OBJ index forms) load to the arrays they were written from.
* The placement transforms meet their closed forms: fitToBBox centres the mesh
  in the box with its longest relative extent filling it; putOnPlane puts the
  lowest vertex along the normal on the plane.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESH_MAIN = os.path.join(ROOT, "tests", "cpp", "mesh_main")
ASSETS = "/root/reference/assets"
PLACE = [-10, -10, -10, 10, 10, 10, 0.8, 10.0, 0.0, 1.0, 0.0, 5.0, 0.0, -10.0]  # appmain.cc bunny placement


load = P.load_mesh


def facade_triangles(path, out, place=None):
    args = [MESH_MAIN, str(path), str(out)] + ([repr(float(a)) for a in place] if place else [])
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return np.fromfile(out, np.float64).reshape(-1, 9)


def ref_triangles(path, place=None):
    cap = 400000
    out = np.zeros((cap, 9), np.float64)
    p = PLACE if place is None else place
    box = np.array(p[0:6], np.float64)
    pn = np.array(p[8:11], np.float64)
    mv = np.array(p[11:14], np.float64)
    n = O.ref().ref_trimesh_place(str(path).encode(), 1 if place else 0, O._ptr(box), p[6], p[7], O._ptr(pn),
                                  O._ptr(mv), O._ptr(out), cap)
    assert n >= 0
    return out[:n]


@pytest.mark.skipif(not (os.path.isdir(ASSETS) and O.ref_available()), reason="reference assets / build absent")
@pytest.mark.parametrize("name", ["bunny.ply", "rt3.ply"])
@pytest.mark.parametrize("placed", [False, True])
def test_reference_assets_load_and_place_bit_exact(tmp_path, name, placed):
    path = os.path.join(ASSETS, name)
    place = PLACE if placed else None
    got = facade_triangles(path, tmp_path / "t.bin", place)
    want = ref_triangles(path, place)
    assert got.shape == want.shape and len(got) > 60000
    assert np.array_equal(got, want)


def _write_ply(path, v, faces, fmt="binary_little_endian", extra=False, dtype="float", count_type="uchar",
               index_type="int"):
    v = np.asarray(v)
    fdt = {"float": "f4", "double": "f8"}[dtype]
    idt = {"int": "i4", "uint": "u4", "short": "i2"}[index_type]
    cdt = {"uchar": "u1", "ushort": "u2", "int": "i4"}[count_type]
    end = {"binary_little_endian": "<", "binary_big_endian": ">", "ascii": "<"}[fmt]
    hdr = ["ply", f"format {fmt} 1.0", "comment written by a test", f"element vertex {len(v)}"]
    hdr += [f"property {dtype} {c}" for c in "xyz"]
    if extra:
        hdr += ["property float nx", "property float ny", "property float nz", "property uchar red"]
    hdr += [f"element face {len(faces)}", f"property list {count_type} {index_type} vertex_indices"]
    if extra:
        hdr += ["property int flags"]
    hdr += ["end_header"]
    with open(path, "wb") as fp:
        fp.write(("\n".join(hdr) + "\n").encode())
        for p in v:
            if fmt == "ascii":
                row = " ".join(repr(float(x)) for x in p) + (" 0 1 0 200" if extra else "")
                fp.write((row + "\n").encode())
            else:
                fp.write(np.asarray(p, end + fdt).tobytes())
                if extra:
                    fp.write(np.asarray([0, 1, 0], end + "f4").tobytes() + bytes([200]))
        for f in faces:
            if fmt == "ascii":
                fp.write((" ".join(str(x) for x in [len(f)] + list(f)) + (" 7" if extra else "") + "\n").encode())
            else:
                fp.write(np.asarray([len(f)], end + cdt).tobytes() + np.asarray(f, end + idt).tobytes())
                if extra:
                    fp.write(np.asarray([7], end + "i4").tobytes())


def _fan(faces):
    out = []
    for f in faces:
        for j in range(1, len(f) - 1):
            out.append([f[0], f[j], f[j + 1]])
    return np.array(out, np.int32)


CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                  np.float32) * np.float32(1.25) + np.float32([0.1, -0.2, 0.3])
CUBE_Q = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [0, 4, 7, 3], [1, 2, 6, 5]]


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
@pytest.mark.parametrize("extra", [False, True])
def test_ply_layouts_load(tmp_path, fmt, extra):
    tris = _fan(CUBE_Q)
    path = tmp_path / "m.ply"
    _write_ply(path, CUBE_V, tris, fmt=fmt, extra=extra)
    v, f = load(path)
    np.testing.assert_array_equal(v, CUBE_V)
    np.testing.assert_array_equal(f, tris)


def test_ply_polygons_double_and_other_list_types(tmp_path):
    path = tmp_path / "q.ply"
    _write_ply(path, CUBE_V.astype(np.float64), CUBE_Q, dtype="double", count_type="ushort", index_type="uint")
    v, f = load(path)
    np.testing.assert_array_equal(v, CUBE_V)
    np.testing.assert_array_equal(f, _fan(CUBE_Q))  # quads fan-triangulated
    assert len(f) == 12


def test_obj_index_forms(tmp_path):
    path = tmp_path / "m.obj"
    lines = ["# cube", "o cube"] + [f"v {x} {y} {z}" for x, y, z in CUBE_V.astype(np.float64)]
    lines += ["vn 0 0 1", "vt 0 0", "s off"]
    for k, q in enumerate(CUBE_Q):
        if k % 3 == 0:
            toks = [str(i + 1) for i in q]                    # 1-based
        elif k % 3 == 1:
            toks = [f"{i + 1}/1/1" for i in q]                 # with texture / normal refs
        else:
            toks = [str(i - len(CUBE_V)) for i in q]          # negative, relative to the end
        lines.append("f " + " ".join(toks))
    path.write_text("\n".join(lines) + "\n")
    v, f = load(path)
    np.testing.assert_array_equal(v, CUBE_V)
    np.testing.assert_array_equal(f, _fan(CUBE_Q))


def test_bad_files_are_rejected(tmp_path):
    p = tmp_path / "bad.ply"
    p.write_bytes(b"not a ply\n")
    with pytest.raises(P.PPPMError):
        load(p)
    p2 = tmp_path / "oob.ply"
    _write_ply(p2, CUBE_V, [[0, 1, 99]])
    with pytest.raises(P.PPPMError):
        load(p2)
    with pytest.raises(P.PPPMError):
        load(tmp_path / "missing.ply")
    p3 = tmp_path / "m.stl"
    p3.write_bytes(b"solid x\n")
    with pytest.raises(P.PPPMError):
        load(p3)


def test_placement_closed_forms(tmp_path):
    rng = np.random.default_rng(4)
    v = rng.normal(size=(500, 3)).astype(np.float32) * np.float32([3.0, 1.0, 0.5]) + np.float32([4, -2, 1])
    f = rng.integers(0, 500, size=(800, 3)).astype(np.int32)
    path = tmp_path / "r.ply"
    _write_ply(path, v, f)
    loaded = facade_triangles(path, tmp_path / "a.bin")
    np.testing.assert_array_equal(loaded.reshape(-1, 3, 3), v.astype(np.float64)[f])
    box = np.array([-10, -10, -10, 10, 10, 10.0])
    plane_d, n = 10.0, np.array([0.0, 1.0, 0.0])
    placed = facade_triangles(path, tmp_path / "b.bin", list(box) + [0.8, plane_d] + list(n) + [5.0, 0.0, -10.0])
    pts = placed.reshape(-1, 3)
    # fitToBBox: the x extent (the largest relative to the cube box) spans 20, then scale 0.8
    used = np.unique(f)
    ext = np.ptp(v.astype(np.float64)[used], axis=0)
    s = 20.0 / ext.max() * 0.8
    np.testing.assert_allclose(np.ptp(pts, axis=0), ext * s, rtol=1e-12)
    # putOnPlane: min over n . x equals -d; translate moves x and z by (5, -10)
    assert abs(pts[:, 1].min() + plane_d) < 1e-12
    c = (pts.min(0) + pts.max(0)) / 2
    assert abs(c[0] - 5.0) < 1e-9 and abs(c[2] + 10.0) < 1e-9
