"""CPU-side checks of the product library: it loads, exports every entry point
include/pppm.h declares, generates the benchmark scenes, and fails loudly
(status codes, no abort) without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "pppm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pppm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(P.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(P.EXPORTED)


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    lib = P.lib()
    rc = lib.pppm_create(C.byref(h), 0)
    assert rc == P.ERR_HIP and not h.value
    assert b"device" in lib.pppm_last_error(None)


def test_radius_schedule_matches_oracle_bitwise():
    for lit in (False, True):
        a = P.radius_schedule(0.0866, 0.7, 256, literal=lit)
        b = O.radius_schedule(0.0866, 0.7, 256, literal=lit)
        assert np.array_equal(a, b)


def test_radius_schedule_rejects_bad_arguments():
    lib = P.lib()
    out = np.zeros(4)
    assert lib.pppm_radius_schedule(-1.0, 0.7, 4, 0, out.ctypes.data_as(C.c_void_p)) == P.ERR_INVALID
    assert lib.pppm_radius_schedule(1.0, 0.0, 4, 0, out.ctypes.data_as(C.c_void_p)) == P.ERR_INVALID


def _edge_counts(tri):
    v = tri.reshape(-1, 3, 3)
    edges = {}
    for t in v:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            k = tuple(sorted((tuple(t[a]), tuple(t[b]))))
            edges[k] = edges.get(k, 0) + 1
    return edges


@pytest.mark.parametrize("f", [1, 2, 8])
def test_icosphere_counts_and_watertight(f):
    s = P.scenegen(P.SCENE_BOX, 1, f, 16, 16)
    sphere_ids = [i for i, b in enumerate(s.bsdfs) if b.type == P.LAMBERTIAN][6:]
    assert len(sphere_ids) == 2
    for sid in sphere_ids:
        tri = s.tri[s.bsdf_id == sid]
        assert len(tri) == 20 * f * f
        # closed 2-manifold: every edge shared by exactly two triangles (bitwise vertices)
        assert set(_edge_counts(tri).values()) == {2}


def test_scene_configs():
    s1 = P.scenegen(P.SCENE_BOX, 1, 0, 256, 256)
    assert s1.n_tris == 12 + 2 + 2 * 1280
    assert len(s1.lights) == 1 and s1.camera.width == 256
    rho = np.array([b.reflectance[:] for b in s1.bsdfs if b.type == P.LAMBERTIAN])
    assert np.all(rho >= 0.2) and np.all(rho <= 0.8)
    s4 = P.scenegen(P.SCENE_TWO_ROOMS, 1, 10, 64, 64)
    assert s4.n_tris == 10 * 2 + 3 * 2 + 2 + 20 * 100
    # deterministic in the seed
    a = P.scenegen(P.SCENE_BOX, 7, 4, 8, 8)
    b = P.scenegen(P.SCENE_BOX, 7, 4, 8, 8)
    c = P.scenegen(P.SCENE_BOX, 8, 4, 8, 8)
    assert np.array_equal(a.tri, b.tri)
    assert not np.array_equal([x.reflectance[:] for x in a.bsdfs], [x.reflectance[:] for x in c.bsdfs])


def test_light_faces_down_into_the_box():
    s = P.scenegen(P.SCENE_BOX, 1, 2, 8, 8)
    L = s.lights[0]
    n = np.cross(L.edge_u[:], L.edge_v[:])
    assert n[1] < 0 and abs(n[0]) < 1e-9 and abs(n[2]) < 1e-9
    em = [i for i, b in enumerate(s.bsdfs) if b.type == P.EMITTER]
    tri = s.tri[s.bsdf_id == em[0]].reshape(-1, 3, 3)
    for t in tri:
        assert np.cross(t[1] - t[0], t[2] - t[0])[1] < 0
