"""ctypes binding of the oracle (liboracle.so) and of the compiled reference
(_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_P = C.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PassParams(C.Structure):
    _fields_ = [("photons", C.c_int64), ("pass_index", C.c_int32), ("max_bounces", C.c_int32),
                ("radius", C.c_double), ("ray_eps", C.c_double), ("seed", C.c_uint64),
                ("scene_id", C.c_uint32), ("pad", C.c_int32)]


class Proposal(C.Structure):
    _fields_ = [("cdf", C.c_void_p), ("lam", C.c_double), ("side", C.c_int32), ("pad", C.c_int32)]


class RenderOpts(C.Structure):
    _fields_ = [("prop", C.c_void_p), ("sss_pts", C.c_void_p), ("sss_nrm", C.c_void_p), ("sss_n", C.c_int64),
                ("sss_area", C.c_double)]


def proposal(cdf, side, lam):
    """orc_proposal over a [side^2][1024] float32 CDF array (kept alive on the struct)."""
    cdf = np.ascontiguousarray(cdf, np.float32).reshape(side * side, 1024)
    pr = Proposal(cdf=cdf.ctypes.data, lam=lam, side=side)
    pr._keep = cdf
    return pr


_ODECL = {
    "orc_philox4x32_10": (None, [_P, _P, _P]),
    "orc_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _P]),
    "orc_scene_create": (_P, [_P, _P, C.c_int64, _P, C.c_int32, _P, C.c_int32]),
    "orc_scene_destroy": (None, [_P]),
    "orc_intersect_batch": (None, [_P, _P, C.c_int64, C.c_int32, _P, _P]),
    "orc_qbvh_stats": (C.c_int32, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "orc_hash": (C.c_uint32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "orc_hashgrid_points": (C.c_int32, [_P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P, _P]),
    "orc_hashgrid_boxes": (C.c_int64, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, C.c_int64, _P, _P]),
    "orc_radius_schedule": (None, [C.c_double, C.c_double, C.c_int32, C.c_int32, _P]),
    "orc_trace_photons": (C.c_int64, [_P, C.POINTER(PassParams), _P, _P, _P, _P, C.c_int64]),
    "orc_trace_photons_prop": (C.c_int64, [_P, C.POINTER(PassParams), C.POINTER(Proposal), _P, _P, _P, _P,
                                           C.c_int64]),
    "orc_emission": (None, [C.POINTER(Proposal), C.POINTER(PassParams), C.c_int64, _P, _P, _P, _P, _P]),
    "orc_dipole_fdr": (C.c_double, [C.c_double]),
    "orc_dipole": (None, [_P, _P, C.c_double, C.c_double, C.c_int32, _P]),
    "orc_dipole_sum": (None, [_P, _P, C.c_double, _P, _P, C.c_int64, C.c_double, _P, C.c_int64, _P]),
    "orc_dipole_octree": (None, [_P, _P, C.c_double, _P, _P, C.c_int64, C.c_double, _P, C.c_int64, C.c_double,
                                 C.c_int32, _P, _P]),
    "orc_eye_pass": (None, [_P, _P, C.POINTER(PassParams), _P, _P, _P, _P, _P]),
    "orc_gather": (None, [_P, _P, _P, C.c_int64, _P, _P, _P, C.c_int64, C.c_double, C.c_int32, _P, _P, _P]),
    "orc_render": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double,
                               C.c_double, C.c_uint64, C.c_uint32, C.c_int32, _P, _P]),
    "orc_render_ex": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, C.c_uint32, C.c_int32, _P, _P, _P]),
    "orc_trace_irradiance_photons": (C.c_int64, [_P, C.POINTER(PassParams), _P, _P, _P, _P, C.c_int64]),
    "orc_scene_set_envmap": (C.c_int32, [_P, _P, C.c_int32, C.c_int32]),
    "orc_envmap_lookup": (None, [_P, _P, C.c_int64, _P]),
    "orc_envmap_cells": (None, [_P, _P, _P]),
    "orc_envmap_sample_dir": (None, [_P, _P, C.c_int64, _P, _P, _P]),
    "orc_render_hachisuka": (C.c_int32, [_P, _P, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_uint64,
                                         C.c_uint32, C.c_int32, _P, _P]),
    "orc_pathtrace": (None, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_uint32, _P]),
    "orc_num_threads": (C.c_int32, []),
    "orc_on_hemisphere": (None, [_P, C.c_double, C.c_double, _P]),
    "orc_total_reflection": (C.c_int32, [C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "orc_bsdf_sample": (None, [C.c_int32, _P, _P, _P, C.c_double, C.c_double, _P, _P]),
    "orc_camera_ray": (None, [_P, C.c_double, C.c_double, _P, _P]),
    "orc_tri_intersect": (C.c_int32, [_P, _P, _P, _P]),
}

_RDECL = {
    "ref_hashgrid_point_buckets": (C.c_int64, [_P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P, C.c_int64, _P]),
    "ref_hashgrid_box_buckets": (C.c_int64, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P,
                                             C.c_int64, _P, C.c_int64, _P]),
    "ref_qbvh_intersect": (None, [_P, C.c_int64, _P, C.c_int64, _P, _P]),
    "ref_triangle_intersect": (C.c_int32, [_P, _P, _P, _P, _P]),
    "ref_on_hemisphere": (None, [_P, C.c_double, C.c_double, _P]),
    "ref_total_reflection": (C.c_int32, [C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "ref_bsdf_sample": (None, [C.c_int32, _P, _P, _P, C.c_double, C.c_double, _P, _P]),
    "ref_camera_ray": (None, [_P, _P, _P, C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double, _P, _P]),
    "ref_dipole": (None, [_P, _P, C.c_double, C.c_double, _P, _P]),
    "ref_dipole_octree": (None, [_P, _P, C.c_double, _P, _P, C.c_int64, C.c_double, _P, C.c_int64, C.c_double, _P]),
    "ref_image_save": (None, [C.c_char_p, _P, C.c_int32, C.c_int32]),
    "ref_image_load": (C.c_int64, [C.c_char_p, _P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ref_envmap_sample": (None, [C.c_char_p, _P, C.c_int64, C.c_int32, _P, _P, _P]),
    "ref_envmap_lookup": (None, [C.c_char_p, _P, C.c_int64, _P]),
    "ref_trimesh_place": (C.c_int64, [C.c_char_p, C.c_int32, _P, C.c_double, C.c_double, _P, _P, _P, C.c_int64]),
}


def _load(path, decl):
    lib = C.CDLL(path)
    for name, (res, args) in decl.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        _orc = _load(ORACLE_PATH, _ODECL)
    return _orc


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_PATH, _RDECL)
    return _ref


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    orc().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def uniforms(seed, kind, pass_index, index, scene, n):
    out = np.zeros(n, np.float64)
    orc().orc_uniforms(seed, kind, pass_index, index, scene, n, _ptr(out))
    return out


class OracleScene:
    """The oracle's copy of a scene (FP64 triangles + restated QBVH)."""

    def __init__(self, scene):
        self.scene = scene
        self._tri = np.ascontiguousarray(scene.tri, np.float32)
        self._ids = np.ascontiguousarray(scene.bsdf_id, np.int32)
        self._bsdf = scene.bsdf_array()
        self._lights = scene.light_array() if scene.lights else None
        self.h = orc().orc_scene_create(_ptr(self._tri), _ptr(self._ids), len(self._ids),
                                        C.cast(self._bsdf, C.c_void_p), len(scene.bsdfs),
                                        C.cast(self._lights, C.c_void_p) if self._lights is not None else None,
                                        len(scene.lights))

    def __del__(self):
        if getattr(self, "h", None):
            orc().orc_scene_destroy(self.h)
            self.h = None

    def set_envmap(self, rgb):
        """Lat-long environment [h][w][3] (row 0 at +y), h, w >= 64."""
        self._env = np.ascontiguousarray(rgb, np.float32)
        rc = orc().orc_scene_set_envmap(self.h, _ptr(self._env), self._env.shape[1], self._env.shape[0])
        assert rc == 0

    def envmap_cells(self):
        pdf, cdf = np.zeros(4096), np.zeros(4096)
        orc().orc_envmap_cells(self.h, _ptr(pdf), _ptr(cdf))
        return pdf, cdf

    def envmap_lookup(self, dirs):
        d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        out = np.zeros_like(d)
        orc().orc_envmap_lookup(self.h, _ptr(d), len(d), _ptr(out))
        return out

    def envmap_sample_dir(self, r):
        r = np.ascontiguousarray(r, np.float64).reshape(-1, 3)
        d, w, c = np.zeros_like(r), np.zeros_like(r), np.zeros(len(r), np.int32)
        orc().orc_envmap_sample_dir(self.h, _ptr(r), len(r), _ptr(d), _ptr(w), _ptr(c))
        return d, w, c

    def intersect(self, rays, brute=False):
        rays = np.ascontiguousarray(rays)
        t = np.zeros(len(rays), np.float32)
        tri = np.zeros(len(rays), np.int32)
        orc().orc_intersect_batch(self.h, _ptr(rays), len(rays), 1 if brute else 0, _ptr(t), _ptr(tri))
        return t, tri

    def qbvh_stats(self):
        n, l, d = C.c_int64(), C.c_int64(), C.c_int32()
        orc().orc_qbvh_stats(self.h, C.byref(n), C.byref(l), C.byref(d))
        return n.value, l.value, d.value

    def _params(self, photons, pass_index, radius, ray_eps, seed, scene_id=0, max_bounces=64):
        return PassParams(photons=photons, pass_index=pass_index, max_bounces=max_bounces,
                          radius=radius, ray_eps=ray_eps, seed=seed, scene_id=scene_id)

    def trace_photons(self, photons, pass_index, ray_eps, seed, scene_id=0, cap=None, prop=None):
        cap = cap or photons * 8 + 1024
        pos = np.zeros((cap, 3), np.float32)
        flux = np.zeros((cap, 3), np.float32)
        nrm = np.zeros((cap, 3), np.float32)
        key = np.zeros(cap, np.uint32)
        p = self._params(photons, pass_index, 0.0, ray_eps, seed, scene_id)
        if prop is None:
            n = orc().orc_trace_photons(self.h, C.byref(p), _ptr(pos), _ptr(flux), _ptr(nrm), _ptr(key), cap)
        else:
            n = orc().orc_trace_photons_prop(self.h, C.byref(p), C.byref(prop), _ptr(pos), _ptr(flux), _ptr(nrm),
                                             _ptr(key), cap)
        assert n <= cap
        self.last_keys = key[:n]
        return pos[:n], flux[:n], nrm[:n]

    def eye_pass(self, cam, pass_index, ray_eps, seed, scene_id=0):
        npix = cam.width * cam.height
        has = np.zeros(npix, np.int32)
        hp = np.zeros((npix, 3), np.float32)
        hn = np.zeros((npix, 3), np.float32)
        hw = np.zeros((npix, 3), np.float32)
        dr = np.zeros((npix, 3), np.float64)
        p = self._params(0, pass_index, 0.0, ray_eps, seed, scene_id)
        orc().orc_eye_pass(self.h, C.byref(cam), C.byref(p), _ptr(has), _ptr(hp), _ptr(hn), _ptr(hw),
                           _ptr(dr))
        return has, hp, hn, hw, dr

    def render(self, cam, photons, passes, r1, alpha=0.7, ray_eps=1e-4, seed=0x7A75, scene_id=0,
               hash_size=0, first_pass=1, prop=None, sss_points=None):
        """P-KZ render.  prop: a Proposal (fixed tables) for the emission;
        sss_points: (pos [n][3], normals [n][3], area per point) of the BSSRDF
        surfaces for the subsurface estimate.  Returns (image, stats) with
        stats = deposits, hit points, irradiance photons, transmitted eye paths."""
        img = np.zeros((cam.height, cam.width, 3), np.float64)
        stats = np.zeros(4, np.int64)
        opts = RenderOpts()
        keep = []
        if prop is not None:
            opts.prop = C.cast(C.pointer(prop), C.c_void_p)
            keep.append(prop)
        if sss_points is not None:
            pts = np.ascontiguousarray(sss_points[0], np.float32).reshape(-1, 3)
            nrm = np.ascontiguousarray(sss_points[1], np.float32).reshape(-1, 3)
            keep += [pts, nrm]
            opts.sss_pts, opts.sss_nrm = pts.ctypes.data, nrm.ctypes.data
            opts.sss_n, opts.sss_area = len(pts), float(sss_points[2])
        rc = orc().orc_render_ex(self.h, C.byref(cam), photons, passes, first_pass, r1, alpha, ray_eps,
                                 seed, scene_id, hash_size, C.byref(opts), _ptr(img), _ptr(stats))
        assert rc == 0
        return img, stats

    def render_hachisuka(self, cam, photons, passes, alpha=0.7, ray_eps=1e-4, seed=0x7A75, scene_id=0, hash_size=0):
        img = np.zeros((cam.height, cam.width, 3), np.float64)
        st = np.zeros(1, np.int64)
        rc = orc().orc_render_hachisuka(self.h, C.byref(cam), photons, passes, alpha, ray_eps, seed, scene_id,
                                        hash_size, _ptr(img), _ptr(st))
        assert rc == 0
        return img, st

    def pathtrace(self, cam, spp, first_sample=0, max_bounces=64, ray_eps=1e-4, seed=0x7A75, scene_id=0):
        """The independent next-event path tracer: mean radiance per pixel."""
        img = np.zeros((cam.height, cam.width, 3), np.float64)
        orc().orc_pathtrace(self.h, C.byref(cam), spp, first_sample, max_bounces, ray_eps, seed, scene_id,
                            _ptr(img))
        return img

    def trace_irradiance_photons(self, photons, pass_index, ray_eps, seed, scene_id=0, cap=None):
        cap = cap or photons * 8 + 1024
        pos = np.zeros((cap, 3), np.float32)
        flux = np.zeros((cap, 3), np.float32)
        nrm = np.zeros((cap, 3), np.float32)
        key = np.zeros(cap, np.uint32)
        p = self._params(photons, pass_index, 0.0, ray_eps, seed, scene_id)
        n = orc().orc_trace_irradiance_photons(self.h, C.byref(p), _ptr(pos), _ptr(flux), _ptr(nrm), _ptr(key), cap)
        assert n <= cap
        return pos[:n], flux[:n], nrm[:n], key[:n]


def hashgrid_points(pos, hash_size, hash_scale, bbox_min):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    sidx = np.zeros(max(n, 1), np.int32)
    cs = np.zeros(hash_size, np.int32)
    ce = np.zeros(hash_size, np.int32)
    bm = np.ascontiguousarray(bbox_min, np.float64)
    orc().orc_hashgrid_points(_ptr(pos), n, hash_size, hash_scale, _ptr(bm), _ptr(sidx), _ptr(cs), _ptr(ce))
    return sidx[:n], cs, ce


def hashgrid_boxes(bmin, bmax, hash_size, hash_scale, bbox_min, cap):
    bmin = np.ascontiguousarray(bmin, np.float64).reshape(-1, 3)
    bmax = np.ascontiguousarray(bmax, np.float64).reshape(-1, 3)
    items = np.zeros(max(cap, 1), np.int32)
    cs = np.zeros(hash_size, np.int32)
    ce = np.zeros(hash_size, np.int32)
    bm = np.ascontiguousarray(bbox_min, np.float64)
    n = orc().orc_hashgrid_boxes(_ptr(bmin), _ptr(bmax), len(bmin), hash_size, hash_scale, _ptr(bm),
                                 _ptr(items), cap, _ptr(cs), _ptr(ce))
    assert n >= 0
    return items[:n], cs, ce


def radius_schedule(r1, alpha, n, literal=False):
    out = np.zeros(n, np.float64)
    orc().orc_radius_schedule(r1, alpha, n, 1 if literal else 0, _ptr(out))
    return out


def gather(ph_pos, ph_flux, ph_nrm, hit_pos, hit_nrm, hit_w, radius, hash_size, bbox_min):
    a = [np.ascontiguousarray(x, np.float32).reshape(-1, 3)
         for x in (ph_pos, ph_flux, ph_nrm, hit_pos, hit_nrm, hit_w)]
    out = np.zeros((max(len(a[3]), 1), 3), np.float64)
    cnt = np.zeros(3, np.int64)
    bm = np.ascontiguousarray(bbox_min, np.float64)
    orc().orc_gather(_ptr(a[0]), _ptr(a[1]), _ptr(a[2]), len(a[0]), _ptr(a[3]), _ptr(a[4]), _ptr(a[5]),
                     len(a[3]), radius, hash_size, _ptr(bm), _ptr(out), _ptr(cnt))
    return out[:len(a[3])], cnt


def emission(prop, photons, pass_index, seed, idx, normal, scene_id=0):
    """Light-sample (u, v), direction, flux factor and cell of photon idx."""
    p = PassParams(photons=photons, pass_index=pass_index, max_bounces=64, radius=0.0, ray_eps=0.0,
                   seed=seed, scene_id=scene_id)
    nl = np.ascontiguousarray(normal, np.float64)
    uv = np.zeros(2)
    d = np.zeros(3)
    w = C.c_double()
    c = C.c_uint32()
    orc().orc_emission(C.byref(prop) if prop is not None else None, C.byref(p), idx, _ptr(nl), _ptr(uv),
                       _ptr(d), C.byref(w), C.byref(c))
    return uv, d, w.value, c.value


def dipole(sigma_a, sigmap_s, eta, d2, fast=False):
    sa = np.ascontiguousarray(sigma_a, np.float64)
    ss = np.ascontiguousarray(sigmap_s, np.float64)
    out = np.zeros(3)
    orc().orc_dipole(_ptr(sa), _ptr(ss), eta, d2, 1 if fast else 0, _ptr(out))
    return out


def dipole_sum(sigma_a, sigmap_s, eta, pts, E, area, query):
    sa = np.ascontiguousarray(sigma_a, np.float64)
    ss = np.ascontiguousarray(sigmap_s, np.float64)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(E, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(query, np.float32).reshape(-1, 3)
    out = np.zeros((len(q), 3))
    orc().orc_dipole_sum(_ptr(sa), _ptr(ss), eta, _ptr(p), _ptr(e), len(p), area, _ptr(q), len(q), _ptr(out))
    return out


def dipole_octree(sigma_a, sigmap_s, eta, pts, E, area, query, max_error=0.05, fast=False):
    """The dipole sum over the octree cut; returns (out [n_q][3], visited [n_q]).
    fast: the reference's exponential in R_d."""
    sa = np.ascontiguousarray(sigma_a, np.float64)
    ss = np.ascontiguousarray(sigmap_s, np.float64)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(E, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(query, np.float32).reshape(-1, 3)
    out = np.zeros((len(q), 3))
    vis = np.zeros(max(len(q), 1), np.int64)
    orc().orc_dipole_octree(_ptr(sa), _ptr(ss), eta, _ptr(p), _ptr(e), len(p), area, _ptr(q), len(q), max_error,
                            1 if fast else 0, _ptr(out), _ptr(vis))
    return out, vis[:len(q)]
