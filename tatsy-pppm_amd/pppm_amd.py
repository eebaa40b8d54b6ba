"""ctypes binding of libpppm_hip.so (the C ABI in include/pppm.h).

Thin plumbing for Python callers (tests, bench.py): numpy arrays in and out,
no compute here.  The library is loaded from this directory; if it is absent
the import fails loudly -- there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpppm_hip.so")

OK, ERR_INVALID, ERR_HIP, ERR_STATE, ERR_CAPACITY, ERR_NOMEM = range(6)
LAMBERTIAN, SPECULAR, PHONG, REFRACTION, BSSRDF, EMITTER = 0x01, 0x02, 0x04, 0x08, 0x10, 0x40
SCENE_BOX, SCENE_BOX_CAUSTIC, SCENE_RANDOM, SCENE_TWO_ROOMS, SCENE_TWO_ROOMS_SSS = 1, 2, 3, 4, 5
GATHER_AUTO, GATHER_LANES, GATHER_TILED, GATHER_WAVE, GATHER_BLOCKS = 0, 1, 2, 3, 4
BVH_DEFAULT, BVH_LBVH, BVH_PLOC = 0, 1, 2
METHOD_PROBABILISTIC, METHOD_HACHISUKA = 0, 1
PROPOSAL_CELLS = 1024
MLP_PARAMS = 76992


class Bsdf(C.Structure):
    _fields_ = [("type", C.c_int32), ("reflectance", C.c_float * 3), ("emission", C.c_float * 3),
                ("ior", C.c_float), ("sigma_a", C.c_float * 3), ("sigmap_s", C.c_float * 3),
                ("eta", C.c_float), ("pad", C.c_float)]


class Light(C.Structure):
    _fields_ = [("corner", C.c_float * 3), ("edge_u", C.c_float * 3), ("edge_v", C.c_float * 3),
                ("emission", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("direction", C.c_double * 3), ("up", C.c_double * 3),
                ("fov_y_deg", C.c_double), ("width", C.c_int32), ("height", C.c_int32)]


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("leaves", C.c_int64), ("depth", C.c_int32),
                ("max_leaf", C.c_int32), ("build_ms", C.c_double), ("builder", C.c_int32),
                ("reserved", C.c_int32)]


class RenderParams(C.Structure):
    _fields_ = [("photons", C.c_int64), ("passes", C.c_int32), ("max_bounces", C.c_int32),
                ("alpha", C.c_double), ("radius1", C.c_double), ("ray_eps", C.c_double),
                ("seed", C.c_uint64), ("scene_id", C.c_uint32), ("hash_size", C.c_int32),
                ("capacity_factor", C.c_double), ("gather_mode", C.c_int32), ("method", C.c_int32)]


class PassStats(C.Structure):
    _fields_ = [("pass_index", C.c_int32), ("hash_size", C.c_int32), ("radius", C.c_double),
                ("photons_emitted", C.c_int64), ("photons_stored", C.c_int64),
                ("photon_segments", C.c_int64), ("eye_segments", C.c_int64),
                ("hitpoints", C.c_int64), ("gather_buckets", C.c_int64),
                ("gather_candidates", C.c_int64), ("gather_inradius", C.c_int64),
                ("photon_ms", C.c_double), ("order_ms", C.c_double), ("grid_ms", C.c_double),
                ("eye_ms", C.c_double), ("gather_ms", C.c_double), ("total_ms", C.c_double),
                ("hitsort_ms", C.c_double), ("proposal_ms", C.c_double), ("train_loss", C.c_double),
                ("proposal_records", C.c_int64), ("proposal_active", C.c_int32), ("retrained", C.c_int32),
                ("irradiance_ms", C.c_double), ("dipole_ms", C.c_double), ("irradiance_photons", C.c_int64),
                ("sss_hits", C.c_int64), ("sss_points", C.c_int64),
                ("gather_tested", C.c_int64), ("gather_blocks", C.c_int64),
                ("photon_nodes", C.c_int64), ("photon_tris", C.c_int64), ("eye_nodes", C.c_int64),
                ("eye_tris", C.c_int64), ("gather_staged", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ProposalParams(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("learning_rate", C.c_double), ("ring_capacity", C.c_int64),
                ("anchors", C.c_int32), ("train_interval", C.c_int32), ("train_steps", C.c_int32),
                ("batch", C.c_int32), ("seed", C.c_uint64)]


class Mesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("vertices", C.POINTER(C.c_float)), ("n_faces", C.c_int64),
                ("faces", C.POINTER(C.c_int32))]


class BvhNode(C.Structure):
    _fields_ = [("lox", C.c_float * 4), ("hix", C.c_float * 4), ("loy", C.c_float * 4), ("hiy", C.c_float * 4),
                ("loz", C.c_float * 4), ("hiz", C.c_float * 4), ("child", C.c_int32 * 4), ("pad", C.c_int32 * 4)]


class SceneDesc(C.Structure):
    _fields_ = [("n_tris", C.c_int64), ("tri_xyz", C.POINTER(C.c_float)),
                ("bsdf_id", C.POINTER(C.c_int32)), ("n_bsdf", C.c_int32),
                ("bsdfs", C.POINTER(Bsdf)), ("n_lights", C.c_int32),
                ("lights", C.POINTER(Light)), ("camera", Camera),
                ("bbox_min", C.c_double * 3), ("bbox_max", C.c_double * 3)]


RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("tri", "<i4"), ("u", "<f4"), ("v", "<f4")])

_P = C.c_void_p
_DECLS = {
    "pppm_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32]),
    "pppm_destroy": (C.c_int, [_P]),
    "pppm_last_error": (C.c_char_p, [_P]),
    "pppm_device_info": (C.c_int, [_P, C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "pppm_scene_upload": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(Bsdf), C.c_int32, C.POINTER(Light), C.c_int32]),
    "pppm_bvh_build": (C.c_int, [_P, C.POINTER(BvhInfo)]),
    "pppm_bvh_build_with": (C.c_int, [_P, C.c_int32, C.POINTER(BvhInfo)]),
    "pppm_intersect": (C.c_int, [_P, _P, C.c_int64, _P]),
    "pppm_bvh_download": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32)]),
    "pppm_mesh_load": (C.c_int, [C.c_char_p, C.POINTER(Mesh), C.c_char_p, C.c_int32]),
    "pppm_mesh_free": (None, [C.POINTER(Mesh)]),
    "pppm_envmap_set": (C.c_int, [_P, _P, C.c_int32, C.c_int32]),
    "pppm_envmap_importance": (C.c_int, [_P, C.c_int32, C.c_int32, _P, _P]),
    "pppm_image_save": (C.c_int, [_P, C.c_char_p, C.c_double]),
    "pppm_image_write": (C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_double]),
    "pppm_image_load": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.c_char_p, C.c_int32]),
    "pppm_image_free": (None, [C.POINTER(C.c_double)]),
    "pppm_hashgrid_build": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P, _P]),
    "pppm_hashgrid_build_noalloc": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P, _P, _P]),
    "pppm_hashgrid_build_boxes": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32, C.c_double, _P, _P,
                                            C.c_int64, C.POINTER(C.c_int64), _P, _P]),
    "pppm_radius_schedule": (C.c_int, [C.c_double, C.c_double, C.c_int32, C.c_int32, _P]),
    "pppm_camera_set": (C.c_int, [_P, C.POINTER(Camera)]),
    "pppm_render_begin": (C.c_int, [_P, C.POINTER(RenderParams)]),
    "pppm_render_pass": (C.c_int, [_P, C.POINTER(PassStats)]),
    "pppm_render_passes": (C.c_int, [_P, C.c_int32, _P]),
    "pppm_image_read": (C.c_int, [_P, _P, C.c_int64]),
    "pppm_count_traversal": (C.c_int, [_P, C.c_int32]),
    "pppm_last_photons": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pppm_last_grid": (C.c_int, [_P, _P, _P, _P, C.c_int32, _P, C.POINTER(C.c_double)]),
    "pppm_last_hitpoints": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pppm_gather": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P, _P, C.c_int64, C.c_double,
                              C.c_int32, _P, C.c_int32, _P, C.POINTER(PassStats)]),
    "pppm_stream_copy": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_double)]),
    "pppm_kahan_probe": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, _P]),
    "pppm_issue_probe": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pppm_subsurface_configure": (C.c_int, [_P, C.c_int64]),
    "pppm_subsurface_cut": (C.c_int, [_P, C.c_double]),
    "pppm_subsurface_points": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "pppm_subsurface_irradiance": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pppm_last_irradiance_photons": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pppm_dipole_eval": (C.c_int, [_P, C.POINTER(Bsdf), _P, _P, C.c_int64, C.c_double, _P, C.c_int64, _P]),
    "pppm_proposal_enable": (C.c_int, [_P, C.POINTER(ProposalParams)]),
    "pppm_proposal_set_table": (C.c_int, [_P, _P, C.c_int32, C.c_double]),
    "pppm_proposal_get_table": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int32)]),
    "pppm_proposal_records": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pppm_mlp_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "pppm_mlp_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "pppm_mlp_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "pppm_mlp_train_step": (C.c_int, [_P, _P, C.c_int64, C.c_double, C.c_int32, _P, C.POINTER(C.c_double)]),
    "pppm_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32]),
    "pppm_batch_last_error": (C.c_char_p, [_P]),
    "pppm_batch_destroy": (C.c_int, [_P]),
    "pppm_batch_add": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(Bsdf), C.c_int32, C.POINTER(Light), C.c_int32,
                                 C.POINTER(Camera), C.POINTER(RenderParams), C.POINTER(C.c_int32)]),
    "pppm_batch_tune": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "pppm_batch_run": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "pppm_batch_image": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "pppm_batch_scene_info": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pppm_batch_scene_trace_keep": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "pppm_scenegen": (C.c_int, [C.c_int32, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(SceneDesc)]),
    "pppm_scenegen_free": (None, [C.POINTER(SceneDesc)]),
}
EXPORTED = tuple(_DECLS)


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run __graft_entry__.build() (no fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in _DECLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PPPMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pppm error {code}: {msg}")
        self.code = code


class Scene:
    """A procedural scene as numpy arrays (copied out of the library)."""

    def __init__(self, tri, bsdf_id, bsdfs, lights, camera, bbox_min, bbox_max):
        self.tri = tri
        self.bsdf_id = bsdf_id
        self.bsdfs = bsdfs
        self.lights = lights
        self.camera = camera
        self.bbox_min = bbox_min
        self.bbox_max = bbox_max

    @property
    def n_tris(self):
        return len(self.bsdf_id)

    def diagonal(self):
        return float(np.linalg.norm(np.asarray(self.bbox_max) - np.asarray(self.bbox_min)))

    def bsdf_array(self):
        return (Bsdf * len(self.bsdfs))(*self.bsdfs)

    def light_array(self):
        return (Light * len(self.lights))(*self.lights)


def scenegen(kind, seed=1, frequency=0, width=256, height=256):
    d = SceneDesc()
    rc = lib().pppm_scenegen(kind, seed, frequency, width, height, C.byref(d))
    if rc != OK:
        raise PPPMError(rc, "scenegen failed")
    try:
        n = d.n_tris
        tri = np.ctypeslib.as_array(d.tri_xyz, shape=(n, 9)).copy() if n else np.zeros((0, 9), np.float32)
        ids = np.ctypeslib.as_array(d.bsdf_id, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        bsdfs = [Bsdf.from_buffer_copy(d.bsdfs[i]) for i in range(d.n_bsdf)]
        lights = [Light.from_buffer_copy(d.lights[i]) for i in range(d.n_lights)]
        cam = Camera.from_buffer_copy(d.camera)
        s = Scene(tri, ids, bsdfs, lights, cam, tuple(d.bbox_min), tuple(d.bbox_max))
    finally:
        lib().pppm_scenegen_free(C.byref(d))
    return s


def load_mesh(path):
    """Vertices [n][3] float32 and triangles [m][3] int32 of a .ply / .obj file."""
    m = Mesh()
    err = C.create_string_buffer(512)
    rc = lib().pppm_mesh_load(str(path).encode(), C.byref(m), err, 512)
    if rc != OK:
        raise PPPMError(rc, err.value.decode())
    try:
        v = np.ctypeslib.as_array(m.vertices, shape=(m.n_vertices, 3)).copy() if m.n_vertices else np.zeros((0, 3), np.float32)
        f = np.ctypeslib.as_array(m.faces, shape=(m.n_faces, 3)).copy() if m.n_faces else np.zeros((0, 3), np.int32)
    finally:
        lib().pppm_mesh_free(C.byref(m))
    return v, f


def envmap_importance(rgb):
    """pdf [4096] and running sums cdf [4096] of an environment map's importance cells."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    pdf, cdf = np.zeros(4096), np.zeros(4096)
    rc = lib().pppm_envmap_importance(_ptr(rgb), rgb.shape[1], rgb.shape[0], _ptr(pdf), _ptr(cdf))
    if rc != OK:
        raise PPPMError(rc, "bad environment map")
    return pdf, cdf


def image_write(path, rgb, gamma=0.0):
    """Write an FP64 [h][w][3] image (top row first) by extension: .png / .bmp
    (8-bit after c^(1/gamma)), .hdr (RGBE), .pfm (FP32)."""
    rgb = np.ascontiguousarray(rgb, np.float64)
    h, w = rgb.shape[:2]
    rc = lib().pppm_image_write(str(path).encode(), _ptr(rgb), w, h, gamma)
    if rc != OK:
        raise PPPMError(rc, f"cannot write {path}")


def image_load(path):
    """FP64 [h][w][3] image of a .hdr, .pfm or .bmp file."""
    p = C.POINTER(C.c_double)()
    w, h = C.c_int32(), C.c_int32()
    err = C.create_string_buffer(512)
    rc = lib().pppm_image_load(str(path).encode(), C.byref(p), C.byref(w), C.byref(h), err, 512)
    if rc != OK:
        raise PPPMError(rc, err.value.decode())
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 3)).copy()
    finally:
        lib().pppm_image_free(p)


def batch_specs(n_scenes, seed):
    """Scene specifications of a batch drawn like the batched-scenes workload:
    triangle counts log-uniform in [1k, 500k], square resolutions from
    {128, 256, 512, 1024}, procedural scenes alternating between the box with
    two icospheres and a random-object scene.  Returns (kind, scene seed,
    icosphere frequency, resolution) per scene; photons per pass are 8 per
    pixel."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_scenes):
        tris = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        res = int(rng.choice([128, 256, 512, 1024]))
        kind = SCENE_RANDOM if k % 2 else SCENE_BOX
        per = 20.0 if kind == SCENE_RANDOM else 40.0   # 20 f^2 triangles per icosphere
        f = max(1, int(round((tris / per) ** 0.5)))
        specs.append((kind, 1000 + k, f, res))
    return specs


def radius_schedule(r1, alpha, n, literal=False):
    out = np.zeros(n, np.float64)
    rc = lib().pppm_radius_schedule(r1, alpha, n, 1 if literal else 0, _ptr(out))
    if rc != OK:
        raise PPPMError(rc, "bad radius schedule arguments")
    return out


class Context:
    """One device context (pppm_create / pppm_destroy)."""

    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.pppm_create(C.byref(h), device)
        if rc != OK:
            raise PPPMError(rc, self.L.pppm_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.L.pppm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != OK:
            raise PPPMError(rc, self.L.pppm_last_error(self.h).decode())

    def device_info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int32()
        hbm = C.c_int64()
        self._chk(self.L.pppm_device_info(self.h, name, 256, C.byref(cu), C.byref(hbm)))
        return {"name": name.value.decode(), "cu": cu.value, "hbm_bytes": hbm.value}

    def stream_copy(self, nbytes):
        ms = C.c_double()
        self._chk(self.L.pppm_stream_copy(self.h, nbytes, C.byref(ms)))
        return ms.value

    def upload(self, scene):
        tri = np.ascontiguousarray(scene.tri, np.float32)
        ids = np.ascontiguousarray(scene.bsdf_id, np.int32)
        lights = scene.light_array() if scene.lights else None
        self._chk(self.L.pppm_scene_upload(self.h, _ptr(tri), _ptr(ids), len(ids), scene.bsdf_array(),
                                           len(scene.bsdfs), lights, len(scene.lights)))

    def build_bvh(self, builder=None):
        """4-wide BVH of the uploaded scene; builder: BVH_LBVH, BVH_PLOC or None (BVH_DEFAULT)."""
        info = BvhInfo()
        if builder is None:
            self._chk(self.L.pppm_bvh_build(self.h, C.byref(info)))
        else:
            self._chk(self.L.pppm_bvh_build_with(self.h, int(builder), C.byref(info)))
        return {f: getattr(info, f) for f, _ in info._fields_ if f != "reserved"}

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        self._chk(self.L.pppm_intersect(self.h, _ptr(rays), len(rays), _ptr(hits)))
        return hits

    def bvh_download(self):
        """Host copy of the built BVH: nodes (structured), leaf-ordered triangles [n][12], root ref."""
        nn, nt, root = C.c_int64(), C.c_int64(), C.c_int32()
        self._chk(self.L.pppm_bvh_download(self.h, None, 0, None, 0, C.byref(nn), C.byref(nt), C.byref(root)))
        nodes = (BvhNode * max(nn.value, 1))()
        tris = np.zeros((max(nt.value, 1), 12), np.float32)
        self._chk(self.L.pppm_bvh_download(self.h, nodes, nn.value, _ptr(tris), nt.value, C.byref(nn), C.byref(nt),
                                           C.byref(root)))
        return nodes, tris[:nt.value], root.value

    def hashgrid(self, pos, hash_size, hash_scale, bbox_min):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = len(pos)
        sidx = np.zeros(max(n, 1), np.int32)
        cs = np.zeros(hash_size, np.int32)
        ce = np.zeros(hash_size, np.int32)
        bm = np.ascontiguousarray(bbox_min, np.float64)
        self._chk(self.L.pppm_hashgrid_build(self.h, _ptr(pos), n, hash_size, hash_scale, _ptr(bm),
                                             _ptr(sidx), _ptr(cs), _ptr(ce)))
        return sidx[:n], cs, ce

    def hashgrid_boxes(self, bmin, bmax, hash_size, hash_scale, bbox_min, cap):
        bmin = np.ascontiguousarray(bmin, np.float64).reshape(-1, 3)
        bmax = np.ascontiguousarray(bmax, np.float64).reshape(-1, 3)
        items = np.zeros(max(cap, 1), np.int32)
        cs = np.zeros(hash_size, np.int32)
        ce = np.zeros(hash_size, np.int32)
        npairs = C.c_int64()
        bm = np.ascontiguousarray(bbox_min, np.float64)
        self._chk(self.L.pppm_hashgrid_build_boxes(self.h, _ptr(bmin), _ptr(bmax), len(bmin), hash_size,
                                                   hash_scale, _ptr(bm), _ptr(items), cap,
                                                   C.byref(npairs), _ptr(cs), _ptr(ce)))
        return items[:npairs.value], cs, ce

    def set_envmap(self, rgb):
        """Lat-long environment map [h][w][3] (None: remove)."""
        if rgb is None:
            self._chk(self.L.pppm_envmap_set(self.h, None, 0, 0))
            return
        self._env = np.ascontiguousarray(rgb, np.float32)
        self._chk(self.L.pppm_envmap_set(self.h, _ptr(self._env), self._env.shape[1], self._env.shape[0]))

    def set_camera(self, cam):
        self._chk(self.L.pppm_camera_set(self.h, C.byref(cam)))
        self.camera = cam

    def begin(self, photons, passes, radius1, alpha=0.7, ray_eps=1e-4, seed=0x7A75, scene_id=0,
              hash_size=0, max_bounces=64, capacity_factor=8.0, gather_mode=GATHER_AUTO, method=METHOD_PROBABILISTIC):
        p = RenderParams(photons=photons, passes=passes, max_bounces=max_bounces, alpha=alpha,
                         radius1=radius1, ray_eps=ray_eps, seed=seed, scene_id=scene_id,
                         hash_size=hash_size, capacity_factor=capacity_factor, gather_mode=gather_mode,
                         method=method)
        self._chk(self.L.pppm_render_begin(self.h, C.byref(p)))
        self.params = p

    def render_pass(self):
        st = PassStats()
        self._chk(self.L.pppm_render_pass(self.h, C.byref(st)))
        return st.as_dict()

    def count_traversal(self, enable=True):
        self._chk(self.L.pppm_count_traversal(self.h, 1 if enable else 0))

    def render_passes(self, count):
        sts = (PassStats * max(count, 1))()
        self._chk(self.L.pppm_render_passes(self.h, count, sts))
        return [sts[k].as_dict() for k in range(count)]

    def image(self):
        w, h = self.camera.width, self.camera.height
        out = np.zeros((h, w, 3), np.float64)
        self._chk(self.L.pppm_image_read(self.h, _ptr(out), out.size))
        return out

    def image_save(self, path, gamma=2.2):
        self._chk(self.L.pppm_image_save(self.h, str(path).encode(), gamma))

    def last_photons(self, cap=None):
        n = C.c_int64()
        self._chk(self.L.pppm_last_photons(self.h, None, None, None, None, 0, C.byref(n)))
        m = n.value if cap is None else min(cap, n.value)
        pos = np.zeros((max(m, 1), 3), np.float32)
        flux = np.zeros((max(m, 1), 3), np.float32)
        nrm = np.zeros((max(m, 1), 3), np.float32)
        key = np.zeros(max(m, 1), np.uint32)
        self._chk(self.L.pppm_last_photons(self.h, _ptr(pos), _ptr(flux), _ptr(nrm), _ptr(key), m,
                                           C.byref(n)))
        return pos[:m], flux[:m], nrm[:m], key[:m]

    def last_grid(self, n, hash_size):
        sidx = np.zeros(max(n, 1), np.int32)
        cs = np.zeros(hash_size, np.int32)
        ce = np.zeros(hash_size, np.int32)
        bm = np.zeros(3, np.float64)
        sc = C.c_double()
        self._chk(self.L.pppm_last_grid(self.h, _ptr(sidx), _ptr(cs), _ptr(ce), hash_size, _ptr(bm),
                                        C.byref(sc)))
        return sidx[:n], cs, ce, bm, sc.value

    def last_hitpoints(self):
        n = C.c_int64()
        self._chk(self.L.pppm_last_hitpoints(self.h, None, None, None, None, 0, C.byref(n)))
        m = n.value
        pos = np.zeros((max(m, 1), 3), np.float32)
        nrm = np.zeros((max(m, 1), 3), np.float32)
        w = np.zeros((max(m, 1), 3), np.float32)
        pix = np.zeros(max(m, 1), np.int32)
        self._chk(self.L.pppm_last_hitpoints(self.h, _ptr(pos), _ptr(nrm), _ptr(w), _ptr(pix), m,
                                             C.byref(n)))
        return pos[:m], nrm[:m], w[:m], pix[:m]

    def gather(self, ph_pos, ph_flux, ph_nrm, hit_pos, hit_nrm, hit_w, radius, hash_size, bbox_min,
               mode=0):
        a = [np.ascontiguousarray(x, np.float32).reshape(-1, 3)
             for x in (ph_pos, ph_flux, ph_nrm, hit_pos, hit_nrm, hit_w)]
        out = np.zeros((max(len(a[3]), 1), 3), np.float32)
        st = PassStats()
        bm = np.ascontiguousarray(bbox_min, np.float64)
        self._chk(self.L.pppm_gather(self.h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), len(a[0]), _ptr(a[3]),
                                     _ptr(a[4]), _ptr(a[5]), len(a[3]), radius, hash_size, _ptr(bm),
                                     mode, _ptr(out), C.byref(st)))
        return out[:len(a[3])], st.as_dict()

    # ---- learned emission proposal ----
    def proposal_enable(self, lam=0.5, anchors=8, train_interval=8, train_steps=64, batch=16384,
                        ring_capacity=16 << 20, learning_rate=2e-3, seed=0x5EED):
        p = ProposalParams(lambda_=lam, learning_rate=learning_rate, ring_capacity=ring_capacity,
                           anchors=anchors, train_interval=train_interval, train_steps=train_steps,
                           batch=batch, seed=seed)
        self._chk(self.L.pppm_proposal_enable(self.h, C.byref(p)))

    def proposal_set_table(self, prob, anchors, lam):
        prob = np.ascontiguousarray(prob, np.float32).reshape(anchors * anchors, PROPOSAL_CELLS)
        self._chk(self.L.pppm_proposal_set_table(self.h, _ptr(prob), anchors, lam))

    def proposal_table(self):
        a = C.c_int32()
        self._chk(self.L.pppm_proposal_get_table(self.h, None, 0, C.byref(a)))
        cdf = np.zeros((a.value * a.value, PROPOSAL_CELLS), np.float32)
        self._chk(self.L.pppm_proposal_get_table(self.h, _ptr(cdf), cdf.size, C.byref(a)))
        return cdf, a.value

    def proposal_records(self):
        n = C.c_int64()
        self._chk(self.L.pppm_proposal_records(self.h, None, 0, C.byref(n)))
        rec = np.zeros((max(n.value, 1), 4), np.float32)
        self._chk(self.L.pppm_proposal_records(self.h, _ptr(rec), n.value, C.byref(n)))
        return rec[:n.value]

    def mlp_set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._chk(self.L.pppm_mlp_set_params(self.h, _ptr(w), w.size))

    def mlp_params(self):
        w = np.zeros(MLP_PARAMS, np.float32)
        self._chk(self.L.pppm_mlp_get_params(self.h, _ptr(w), w.size))
        return w

    def mlp_forward(self, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((max(len(uv), 1), PROPOSAL_CELLS), np.float32)
        self._chk(self.L.pppm_mlp_forward(self.h, _ptr(uv), len(uv), _ptr(out)))
        return out[:len(uv)]

    def mlp_train_step(self, rec, learning_rate=0.0, step=1):
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 4)
        grad = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        self._chk(self.L.pppm_mlp_train_step(self.h, _ptr(rec), len(rec), learning_rate, step, _ptr(grad),
                                             C.byref(loss)))
        return grad, loss.value

    def kahan_probe(self, values, lanes=64):
        v = np.ascontiguousarray(values, np.float32)
        s = np.zeros(lanes, np.float32)
        c = np.zeros(lanes, np.float32)
        self._chk(self.L.pppm_kahan_probe(self.h, _ptr(v), len(v), lanes, _ptr(s), _ptr(c)))
        return s, c

    def issue_probe(self):
        v, t, c = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.pppm_issue_probe(self.h, C.byref(v), C.byref(t), C.byref(c)))
        return {"valu_cycles": v.value, "trans_cycles": t.value, "clock_hz": c.value}

    # ---- subsurface scattering ----
    def subsurface_configure(self, points):
        self._chk(self.L.pppm_subsurface_configure(self.h, points))

    def subsurface_cut(self, max_error):
        """0: the exact dipole sum; > 0: the octree cut with this error bound."""
        self._chk(self.L.pppm_subsurface_cut(self.h, max_error))

    def subsurface_points(self):
        n = C.c_int64()
        area = C.c_double()
        self._chk(self.L.pppm_subsurface_points(self.h, None, None, 0, C.byref(n), C.byref(area)))
        pos = np.zeros((max(n.value, 1), 3), np.float32)
        nrm = np.zeros((max(n.value, 1), 3), np.float32)
        self._chk(self.L.pppm_subsurface_points(self.h, _ptr(pos), _ptr(nrm), n.value, C.byref(n), C.byref(area)))
        return pos[:n.value], nrm[:n.value], area.value

    def subsurface_irradiance(self):
        n = C.c_int64()
        self._chk(self.L.pppm_subsurface_irradiance(self.h, None, 0, C.byref(n)))
        E = np.zeros((max(n.value, 1), 3), np.float32)
        self._chk(self.L.pppm_subsurface_irradiance(self.h, _ptr(E), n.value, C.byref(n)))
        return E[:n.value]

    def last_irradiance_photons(self):
        n = C.c_int64()
        self._chk(self.L.pppm_last_irradiance_photons(self.h, None, None, None, None, 0, C.byref(n)))
        m = n.value
        pos = np.zeros((max(m, 1), 3), np.float32)
        flux = np.zeros((max(m, 1), 3), np.float32)
        nrm = np.zeros((max(m, 1), 3), np.float32)
        key = np.zeros(max(m, 1), np.uint32)
        self._chk(self.L.pppm_last_irradiance_photons(self.h, _ptr(pos), _ptr(flux), _ptr(nrm), _ptr(key), m,
                                                      C.byref(n)))
        return pos[:m], flux[:m], nrm[:m], key[:m]

    def dipole_eval(self, material, pts, E, area, query):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        E = np.ascontiguousarray(E, np.float32).reshape(-1, 3)
        q = np.ascontiguousarray(query, np.float32).reshape(-1, 3)
        out = np.zeros((max(len(q), 1), 3), np.float32)
        self._chk(self.L.pppm_dipole_eval(self.h, C.byref(material), _ptr(pts), _ptr(E), len(pts), area, _ptr(q),
                                          len(q), _ptr(out)))
        return out[:len(q)]


class Batch:
    """Many scenes rendered by one hipGraph-captured pass loop (pppm_batch_*)."""

    def __init__(self, device=0, branches=4):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.pppm_batch_create(C.byref(h), device, branches)
        if rc != OK:
            raise PPPMError(rc, self.L.pppm_batch_last_error(None).decode())
        self.h = h
        self.cams = []

    def _chk(self, rc):
        if rc != OK:
            raise PPPMError(rc, self.L.pppm_batch_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.pppm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scene, photons, passes, radius1, hash_size, alpha=0.7, ray_eps=1e-4, seed=0x7A75, scene_id=0,
            max_bounces=64, capacity_factor=4.0, gather_mode=GATHER_LANES):
        tri = np.ascontiguousarray(scene.tri, np.float32)
        ids = np.ascontiguousarray(scene.bsdf_id, np.int32)
        p = RenderParams(photons=photons, passes=passes, max_bounces=max_bounces, alpha=alpha, radius1=radius1,
                         ray_eps=ray_eps, seed=seed, scene_id=scene_id, hash_size=hash_size,
                         capacity_factor=capacity_factor, gather_mode=gather_mode)
        idx = C.c_int32()
        self._chk(self.L.pppm_batch_add(self.h, _ptr(tri), _ptr(ids), len(ids), scene.bsdf_array(), len(scene.bsdfs),
                                        scene.light_array(), len(scene.lights), C.byref(scene.camera), C.byref(p),
                                        C.byref(idx)))
        self.cams.append(scene.camera)
        return idx.value

    def tune(self, gather=True, grid=True):
        self._chk(self.L.pppm_batch_tune(self.h, 1 if gather else 0, 1 if grid else 0))

    def run(self, passes):
        ms = C.c_double()
        self._chk(self.L.pppm_batch_run(self.h, passes, C.byref(ms)))
        return ms.value

    def image(self, index):
        cam = self.cams[index]
        out = np.zeros((cam.height, cam.width, 3), np.float64)
        self._chk(self.L.pppm_batch_image(self.h, index, _ptr(out), out.size))
        return out

    def info(self, index):
        m, g, b = C.c_int32(), C.c_int32(), C.c_int32()
        c = C.c_double()
        self._chk(self.L.pppm_batch_scene_info(self.h, index, C.byref(m), C.byref(g), C.byref(b), C.byref(c)))
        k = C.c_int32()
        self._chk(self.L.pppm_batch_scene_trace_keep(self.h, index, C.byref(k)))
        return {"gather_mode": m.value, "photon_grid": g.value, "branch": b.value, "cost_ms": c.value,
                "trace_keep": k.value}
