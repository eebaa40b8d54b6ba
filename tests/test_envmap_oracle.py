"""The environment-map emitter of the oracle (row f1), pinned:

* against the compiled reference's Envmap (when its build is present) on a
  map written as .hdr and read back by both: sampleFromDir bit for bit; the
  importance cells bit for bit (uniforms placed exactly on, and one ulp
  above, the reference's running sums pick the same cells); the direction
  of every draw bit for bit and L / pdf against the reference's photon flux;
* by a closed form of the physically scaled emission: a diffuse floor under
  an environment receives E = sum over upper-hemisphere texels of
  L (2 pi / w) (sin^2 t1 - sin^2 t0) / 2, so it shows rho E / pi; an eye ray
  that leaves the scene shows the texel it points at.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
import pppm_amd as P
import scenes as S


def env_image(w=256, h=128, seed=3):
    rng = np.random.default_rng(seed)
    img = 0.2 + 0.3 * rng.random((h, w, 3))
    img[20:26, 40:52] = [30.0, 25.0, 20.0]   # a sun in the upper hemisphere
    img[100:110, 200:230] = [0.0, 0.0, 0.0]  # a black patch below the horizon
    return img


def env_scene_oracle(env, parts, bsdfs, cam):
    sc = S.scene(parts, bsdfs, [], cam)
    os_ = O.OracleScene(sc)
    os_.set_envmap(env)
    return sc, os_


@pytest.fixture(scope="module")
def env_hdr(tmp_path_factory):
    path = tmp_path_factory.mktemp("env") / "env.hdr"
    P.image_write(path, env_image())
    return path, P.image_load(path)  # the map as both sides read it


def _oracle_env(env):
    floor = S.quad((-2, 0, -2), (0, 0, 4), (4, 0, 0))
    cam = S.camera((0, 1, 0), (0, -1, 0), (0, 0, 1), 20.0, 8, 8)
    _, os_ = env_scene_oracle(env, [(floor, 0)], [S.bsdf(P.LAMBERTIAN, (0.5, 0.5, 0.5))], cam)
    return os_


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference absent")


@needs_ref
def test_env_lookup_matches_reference(env_hdr):
    path, env = env_hdr
    os_ = _oracle_env(env)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:6] = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1]]
    want = np.zeros_like(d)
    O.ref().ref_envmap_lookup(str(path).encode(), O._ptr(d), len(d), O._ptr(want))
    np.testing.assert_array_equal(os_.envmap_lookup(d), want)


@needs_ref
def test_env_importance_cells_and_draws_match_reference(env_hdr):
    path, env = env_hdr
    os_ = _oracle_env(env)
    rng = np.random.default_rng(6)
    # a uniform grid of cell picks and random ones
    r0 = np.linspace(0.0, 1.0, 4096, endpoint=False)[1:]
    picks = np.concatenate([r0, rng.random(20000)])
    jit = rng.random((len(picks), 2))
    r = np.column_stack([picks, jit])
    d, lpdf, cells = os_.envmap_sample_dir(r)
    N = 100000
    pos, flux, dref = (np.zeros((len(r), 3)) for _ in range(3))
    O.ref().ref_envmap_sample(str(path).encode(), O._ptr(np.ascontiguousarray(r)), len(r), N, O._ptr(pos),
                              O._ptr(flux), O._ptr(dref))
    np.testing.assert_array_equal(-dref, d)                 # the same cell and direction, bit for bit
    np.testing.assert_array_equal(pos, 50.0 * d)
    h, w = env.shape[:2]
    np.testing.assert_allclose(flux, lpdf * (w * h / (4 * math.pi * 64 * 64)) / N, rtol=1e-14)
    assert len(np.unique(cells)) > 300


@needs_ref
def test_env_running_sums_bit_exact(env_hdr):
    """Uniforms equal to the reference's running sums (recovered from its own
    cell picks by bisection on the FP64 line) give the same cells."""
    path, env = env_hdr
    os_ = _oracle_env(env)
    rng = np.random.default_rng(9)
    cells = rng.integers(0, 4096, 64)
    probes = []
    for c in cells:
        # the oracle's sum for cell c: smallest r whose pick is c ... then the
        # reference must pick c at that r and c + 1 (or later) one ulp below
        # the next sum; bisect on the oracle
        lo, hi = 0.0, 1.0
        for _ in range(80):
            mid = (lo + hi) / 2
            _, _, k = os_.envmap_sample_dir([[mid, 0.5, 0.5]])
            if k[0] >= c:
                hi = mid
            else:
                lo = mid
        probes += [hi, np.nextafter(hi, 0.0)]
    r = np.column_stack([probes, np.full(len(probes), 0.5), np.full(len(probes), 0.5)])
    d, _, _ = os_.envmap_sample_dir(r)
    pos, flux, dref = (np.zeros((len(r), 3)) for _ in range(3))
    O.ref().ref_envmap_sample(str(path).encode(), O._ptr(np.ascontiguousarray(r)), len(r), 1000, O._ptr(pos),
                              O._ptr(flux), O._ptr(dref))
    np.testing.assert_array_equal(-dref, d)


def hemisphere_irradiance(env):
    """E on an upward-facing plane from a lat-long map (exact for texels)."""
    h, w = env.shape[:2]
    E = np.zeros(3)
    for iy in range(h // 2):  # theta in [0, pi/2]
        t0, t1 = math.pi * iy / h, math.pi * (iy + 1) / h
        E += env[iy].sum(axis=0) * (2 * math.pi / w) * (math.sin(t1) ** 2 - math.sin(t0) ** 2) / 2
    return E


def test_env_lit_floor_closed_form():
    env = env_image()
    env32 = env.astype(np.float32).astype(np.float64)
    floor = S.quad((-2, 0, -2), (0, 0, 4), (4, 0, 0))
    cam = S.camera((0, 1, 0), (0, -1, 0), (0, 0, 1), 20.0, 16, 16)
    sc, os_ = env_scene_oracle(env, [(floor, 0)], [S.bsdf(P.LAMBERTIAN, (0.5, 0.5, 0.5))], cam)
    img, st = os_.render(cam, 400000, 4, 0.01 * sc.diagonal())
    expect = 0.5 / math.pi * hemisphere_irradiance(env32)
    got = img.reshape(-1, 3).mean(0)
    np.testing.assert_allclose(got / expect, 1.0, atol=0.03)


def test_eye_ray_into_the_environment_shows_its_texel():
    env = env_image()
    floor = S.quad((-2, 0, -2), (0, 0, 4), (4, 0, 0))
    cam = S.camera((0, 1, 0), (0, 1, 0), (0, 0, 1), 40.0, 16, 16)   # looking up: every ray escapes
    sc, os_ = env_scene_oracle(env, [(floor, 0)], [S.bsdf(P.LAMBERTIAN, (0.5, 0.5, 0.5))], cam)
    has, hp, hn, hw, dr = os_.eye_pass(cam, 1, 1e-4, 0x7A75)
    assert not has.any()
    dirs = np.zeros((256, 3))
    for pix in range(256):
        u = O.uniforms(0x7A75, 2, 1, pix, 0, 2)
        o, d = np.zeros(3), np.zeros(3)
        O.orc().orc_camera_ray(O.C.byref(cam), pix % 16 + u[0], (15 - pix // 16) + u[1], O._ptr(o), O._ptr(d))
        dirs[pix] = d
    np.testing.assert_array_equal(dr, os_.envmap_lookup(dirs))


@pytest.mark.parametrize("shape", [(128, 256), (64, 64), (100, 333)])
def test_product_importance_cells_equal_oracle(shape):
    """The library's host-side builder (pppm_envmap_importance) and the
    oracle's restatement give the same cells and sums bit for bit, also for
    maps whose size is not a multiple of 64."""
    env = env_image(w=shape[1], h=shape[0], seed=shape[1])
    os_ = _oracle_env(env)
    pdf, cdf = P.envmap_importance(env)
    opdf, ocdf = os_.envmap_cells()
    np.testing.assert_array_equal(pdf, opdf)
    np.testing.assert_array_equal(cdf, ocdf)
    assert abs(cdf[-1] - 1.0) < 1e-9 and np.all(np.diff(cdf) >= 0)
