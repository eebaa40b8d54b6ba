"""The environment-map emitter on the GPU against the oracle (row f1), and
image files written from the device (row f4), through the C ABI.

Same seeds on both sides, so GPU and oracle draw the same numbers:
* photons of an environment-lit open scene: the same (photon, bounce) keys,
  positions and flux as the oracle's FP64 restatement;
* whole renders (environment alone, and with an area light beside it) within
  5 % of the seed-to-seed noise and below 1e-2 relative RMSE;
* eye rays leaving the scene show their texel;
* a black environment renders like none, bit for bit (the kernels built with
  and without the environment-light code).
"""
import sys

import numpy as np
import pytest

import oracle as O
import pppm_amd as P
import scenes as S

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def env_image(w=256, h=128, seed=3):
    rng = np.random.default_rng(seed)
    img = 0.2 + 0.3 * rng.random((h, w, 3))
    img[20:26, 40:52] = [30.0, 25.0, 20.0]
    img[100:110, 200:230] = [0.0, 0.0, 0.0]
    return img.astype(np.float32)


def open_scene(res=64, with_light=False, light_down=False):
    floor = S.quad((-3, 0, -3), (0, 0, 6), (6, 0, 0))
    ball = S.icosphere(8, 0.7, (0.5, 0.7, 0.0))
    glass = S.icosphere(8, 0.5, (-1.0, 0.5, 0.8))
    bs = [S.bsdf(P.LAMBERTIAN, (0.6, 0.5, 0.4)), S.bsdf(P.LAMBERTIAN, (0.3, 0.6, 0.7)),
          S.bsdf(P.REFRACTION, (0.99, 0.99, 0.99))]
    parts = [(floor, 0), (ball, 1), (glass, 2)]
    lights = []
    if with_light:
        bs.append(S.bsdf(P.EMITTER, (0, 0, 0), emission=(4.0, 4.0, 4.0)))
        if light_down:  # emitting along (1, 0, 0) x (0, 0, 1) = -y, onto the objects
            q, L = S.emitter_quad((-0.5, 2.5, -0.5), (1, 0, 0), (0, 0, 1), (4.0, 4.0, 4.0), 3)
        else:
            q, L = S.emitter_quad((-0.5, 2.5, -0.5), (0, 0, 1), (1, 0, 0), (4.0, 4.0, 4.0), 3)
        parts.append(q)
        lights.append(L)
    cam = S.camera((0, 1.2, 4.5), (0, -0.25, -1), (0, 1, 0), 50.0, res, res)
    return S.scene(parts, bs, lights, cam)


def _setup(ctx, s, env, photons, passes, seed=0x7A75):
    ctx.upload(s)
    ctx.build_bvh()
    ctx.set_envmap(env)
    ctx.set_camera(s.camera)
    ctx.begin(photons, passes, 0.01 * s.diagonal(), seed=seed)


def _oracle(s, env):
    os_ = O.OracleScene(s)
    os_.set_envmap(env)
    return os_


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.mean(b))


def test_env_photons_match_oracle(gpu_ctx):
    say("environment photons vs oracle")
    s, env = open_scene(), env_image()
    N = 100000
    _setup(gpu_ctx, s, env, N, 1)
    st = gpu_ctx.render_pass()
    pos, flux, nrm, key = gpu_ctx.last_photons()
    os_ = _oracle(s, env)
    opos, oflux, onrm = os_.trace_photons(N, 1, 1e-4, 0x7A75)
    okey = os_.last_keys
    assert st["photons_stored"] > 0.3 * N
    common, gi, oi = np.intersect1d(key, okey, return_indices=True)
    assert len(common) >= 0.995 * max(len(key), len(okey))
    dp = np.linalg.norm(pos[gi] - opos[oi], axis=1)
    assert np.mean(dp < 1e-3) > 0.995
    rf = np.abs(flux[gi] - oflux[oi]).max(axis=1) / np.abs(oflux[oi]).max(axis=1)
    assert np.mean(rf < 1e-4) > 0.995
    # total flux: the same estimate of the power the scene receives
    assert abs(flux.sum() / oflux.sum() - 1.0) < 5e-3


@pytest.mark.parametrize("with_light", [False, True])
def test_env_render_matches_oracle(gpu_ctx, with_light):
    say(f"environment render vs oracle (area light too: {with_light})")
    s, env = open_scene(with_light=with_light), env_image()
    N, passes, r1 = 150000, 3, 0.01 * s.diagonal()
    _setup(gpu_ctx, s, env, N, passes)
    for _ in range(passes):
        gpu_ctx.render_pass()
    g = gpu_ctx.image()
    os_ = _oracle(s, env)
    o, _ = os_.render(s.camera, N, passes, r1)
    o2, _ = os_.render(s.camera, N, passes, r1, seed=0x7A76)
    e, noise = rel_rmse(g, o), rel_rmse(o2, o)
    say(f"relative RMSE {e:.3e}, noise {noise:.3e}")
    assert np.all(np.isfinite(g)) and g.min() >= 0.0
    assert e < 0.05 * noise and e < 1e-2


def test_eye_rays_leaving_the_scene_show_their_texel(gpu_ctx):
    say("sky pixels")
    s, env = open_scene(), env_image()
    # beside the objects, looking up: every ray escapes
    cam = S.camera((0, 0.2, 2.5), (0, 1, 0.2), (0, 0, 1), 60.0, 32, 32)
    s.camera = cam
    _setup(gpu_ctx, s, env, 20000, 1)
    gpu_ctx.render_pass()
    g = gpu_ctx.image()
    has, hp, hn, hw, dr = _oracle(s, env).eye_pass(cam, 1, 1e-4, 0x7A75)
    assert not has.any()
    same = np.all(g.reshape(-1, 3) == dr.astype(np.float32).astype(np.float64), axis=1)
    # the FP32 camera ray can cross a texel boundary where the FP64 one does not
    assert same.mean() > 0.99, same.mean()


def test_black_environment_renders_like_none(gpu_ctx):
    """An all-black environment map runs the photon and eye kernels built with
    the environment-light code, but is never chosen to emit (its power is 0)
    and shows black: photons and image must equal, bit for bit, those of the
    same scene without one, which runs the kernels built without that code."""
    say("black environment vs none (kernels with and without the environment code)")
    s = open_scene(with_light=True, light_down=True)
    N, passes = 60000, 3
    out = []
    try:
        for env in (None, np.zeros((64, 128, 3), np.float32)):
            gpu_ctx.upload(s)
            gpu_ctx.build_bvh()
            gpu_ctx.set_envmap(env)
            gpu_ctx.set_camera(s.camera)
            gpu_ctx.begin(N, passes, 0.01 * s.diagonal(), seed=0x7A75)
            gpu_ctx.render_pass()
            photons = gpu_ctx.last_photons()
            gpu_ctx.render_passes(passes - 1)  # pipelined passes too
            out.append((gpu_ctx.image(), photons))
    finally:
        gpu_ctx.set_envmap(None)
    (a, pa), (b, pb) = out
    assert np.any(a > 0) and len(pa[0]) > 0.3 * N
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)
    assert np.array_equal(a, b)


def test_envmap_arguments_are_checked(gpu_ctx):
    s = open_scene()
    gpu_ctx.upload(s)
    with pytest.raises(P.PPPMError):
        gpu_ctx.set_envmap(np.ones((32, 64, 3), np.float32))  # below 64 x 64
    bad = env_image()
    bad[3, 3, 0] = -1.0
    with pytest.raises(P.PPPMError):
        gpu_ctx.set_envmap(bad)
    gpu_ctx.set_envmap(env_image())
    gpu_ctx.set_envmap(None)


def test_image_save_matches_host_writer(gpu_ctx, tmp_path):
    say("device-encoded image files")
    s = P.scenegen(P.SCENE_BOX, 1, 0, 48, 40)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.begin(30000, 2, 0.01 * s.diagonal())
    gpu_ctx.render_pass()
    gpu_ctx.render_pass()
    img = gpu_ctx.image()
    from test_image_io import decode_png
    for ext in ("hdr", "pfm"):  # frexp and the FP32 cast: the same bytes
        gpu_ctx.image_save(tmp_path / f"d.{ext}", 2.2)
        P.image_write(tmp_path / f"h.{ext}", img, 2.2)
        assert (tmp_path / f"d.{ext}").read_bytes() == (tmp_path / f"h.{ext}").read_bytes(), ext
    # 8-bit after c^(1/2.2): the device's and the host's pow may differ in the
    # last bit, which moves a byte only where the product sits on an integer
    gpu_ctx.image_save(tmp_path / "d.png", 2.2)
    P.image_write(tmp_path / "h.png", img, 2.2)
    a, b = decode_png(tmp_path / "d.png").astype(int), decode_png(tmp_path / "h.png").astype(int)
    assert np.abs(a - b).max() <= 1 and np.mean(a != b) < 1e-3
    gpu_ctx.image_save(tmp_path / "d.bmp", 2.2)
    P.image_write(tmp_path / "h.bmp", img, 2.2)
    a, b = P.image_load(tmp_path / "d.bmp"), P.image_load(tmp_path / "h.bmp")
    assert np.abs(a - b).max() <= 1.0 / 255 + 1e-12 and np.mean(a != b) < 1e-3
