"""Pins the oracle's composed passes -- the eye path, the whole render, the
subsurface composition -- against closed forms and an independent estimator.

Each check is chosen so that a plausible mistake in the composition (a
dropped 1/pi, the Russian-roulette weight applied twice or not at all, a
wrong reflection, the cone-filter normalisation, the Fresnel weight of the
transmitted path) fails it:

* sphere furnace: inside a closed sphere of uniform albedo, a cosine emitter
  on its surface and any uniform radiosity both give uniform irradiance, so
  the exitant radiance is L = rho Phi / (pi A (1 - rho)) everywhere;
* mirror corridor: an eye ray between two parallel mirrors reaches the back
  wall where the unfolded straight ray does, after k = floor(z) reflections,
  with weight prod(rho_mirror) rho_wall / pi divided by the roulette
  probabilities of bounces >= 3 (ppm_probabilistic.cc:256-268);
* translucent plane under a parallel emitter: the pixel expectation is
  rho F_r Le + F_t (1 - F_dr) / pi E I_Rd with I_Rd the area integral of
  the dipole profile (quadrature of the pinned R_d) and E = kappa Le;
* the box scene: the photon mapping render agrees with an independent FP64
  path tracer with next-event estimation (a different estimator of the same
  radiance) to within the Monte Carlo noise and the known radius bias.
"""
import math

import numpy as np
import pytest

import oracle as O
import pppm_amd as P
import scenes as S


def f32(v):
    return np.float32(v).astype(np.float64)


# ---- sphere furnace ----------------------------------------------------------
def test_render_sphere_furnace_closed_form():
    rho, le = (0.3, 0.5, 0.7), 4.0
    sc, sph = S.sphere_furnace(rho=rho, le=(le,) * 3)
    A = S.tri_area(sph).sum()
    phi = math.pi * f32(le) * 0.1 * 0.1
    expect = f32(rho) * phi / (math.pi * A * (1.0 - f32(rho)))
    os_ = O.OracleScene(sc)
    img, st = os_.render(sc.camera, 100000, 4, 0.01 * sc.diagonal())
    got = img.reshape(-1, 3).mean(0)
    # every pixel sees the sphere: one hit point per pixel and pass
    assert st[1] == 4 * sc.camera.width * sc.camera.height
    np.testing.assert_allclose(got / expect, 1.0, atol=0.02)
    # the uniformity itself: no image half differs from the other
    im = img.mean(2)
    assert abs(im[:, :16].mean() / im[:, 16:].mean() - 1) < 0.03


def test_pathtracer_sphere_furnace_closed_form():
    rho, le = (0.3, 0.5, 0.7), 4.0
    sc, sph = S.sphere_furnace(rho=rho, le=(le,) * 3)
    A = S.tri_area(sph).sum()
    phi = math.pi * f32(le) * 0.1 * 0.1
    expect = f32(rho) * phi / (math.pi * A * (1.0 - f32(rho)))
    got = O.OracleScene(sc).pathtrace(sc.camera, 64).reshape(-1, 3).mean(0)
    np.testing.assert_allclose(got / expect, 1.0, atol=0.02)


# ---- mirror corridor ---------------------------------------------------------
def _corridor(length=20.0):
    rm, rw, le = (0.9, 0.8, 0.7), (0.6, 0.5, 0.4), (2.0, 3.0, 4.0)
    parts = [(S.quad((-1, -5, 0), (length + 2, 0, 0), (0, 10, 0)), 0),   # mirror z = 0
             (S.quad((-1, -5, 1), (0, 10, 0), (length + 2, 0, 0)), 0),   # mirror z = 1
             (S.quad((length, -5, 0), (0, 0, 1), (0, 5, 0)), 1)]          # Lambertian wall x = L, y < 0
    eq, lt = S.emitter_quad((length, 0, 0), (0, 0, 1), (0, 5, 0), le, 2)  # emitter x = L, y > 0, facing -x
    bs = [S.bsdf(P.SPECULAR, rm), S.bsdf(P.LAMBERTIAN, rw), S.bsdf(P.EMITTER, (0, 0, 0), emission=le)]
    cam = S.camera((0, 0, 0.5), (1, 0, 0), (0, 1, 0), 5.0, 64, 8)
    return S.scene(parts + [eq], bs, [lt], cam), f32(rm), f32(rw), f32(le)


def test_eye_pass_mirror_corridor_closed_form():
    length = 20.0
    sc, rm, rw, le = _corridor(length)
    cam = sc.camera
    W, H = cam.width, cam.height
    os_ = O.OracleScene(sc)
    seed = 0x7A75
    n_hit = n_dir = 0
    surv = {}
    for ps in range(1, 5):
        has, hp, hn, hw, dr = os_.eye_pass(cam, ps, 1e-4, seed)
        for pix in range(W * H):
            x, row = pix % W, pix // W
            u = O.uniforms(seed, 2, ps, pix, 0, 2)  # the eye stream's jitter dimensions
            oo, dd = np.zeros(3), np.zeros(3)
            O.orc().orc_camera_ray(O.C.byref(cam), x + u[0], (H - 1 - row) + u[1], O._ptr(oo), O._ptr(dd))
            # unfold the corridor: the straight ray meets x = L at z_u; each
            # integer crossing of z is one reflection
            t = (length - oo[0]) / dd[0]
            zu, y = oo[2] + t * dd[2], oo[1] + t * dd[1]
            m = math.floor(zu)
            fr = zu - m
            if min(fr, 1 - fr) < 2e-3 or abs(y) < 2e-3:
                continue  # corner or seam: ray offsets decide
            z = fr if m % 2 == 0 else 1 - fr
            k = abs(m)
            rr_mirror = rm.max() ** max(0, k - 3)  # roulettes at mirror bounces 3 .. k-1
            if y < 0:
                surv.setdefault(k, []).append(has[pix])
                if has[pix]:
                    n_hit += 1
                    expect_w = rm ** k * rw / math.pi / (rr_mirror * (rw.max() if k >= 3 else 1.0))
                    np.testing.assert_allclose(hp[pix], (length, y, z), atol=2e-3)
                    np.testing.assert_allclose(hn[pix], (-1, 0, 0), atol=1e-12)
                    np.testing.assert_allclose(hw[pix], expect_w, rtol=1e-5)
                else:
                    assert np.all(hw[pix] == 0)
            else:
                assert not has[pix]
                if np.any(dr[pix] != 0):
                    n_dir += 1
                    np.testing.assert_allclose(dr[pix], rm ** k * le / rr_mirror, rtol=1e-9)
    assert n_hit > 500 and n_dir > 700
    # survival of the eye roulette: prod of max albedos from bounce 3 on
    z2 = 0.0
    for k, v in surv.items():
        p = (rm.max() ** max(0, k - 3)) * (rw.max() if k >= 3 else 1.0)
        v = np.asarray(v, np.float64)
        if p == 1.0:
            assert v.all()
        else:
            z2 += (v.sum() - p * len(v)) ** 2 / (len(v) * p * (1 - p))
    assert z2 < 25.0  # chi-square with ~5 degrees of freedom, far tail


def test_eye_pass_sees_only_the_front_of_an_emitter():
    le = (2.0, 3.0, 4.0)
    eq, lt = S.emitter_quad((5, -1, -1), (0, 0, 2), (0, 2, 0), le, 0)      # facing -x: seen
    bq, lb = S.emitter_quad((-5, -1, -1), (0, 0, 2), (0, 2, 0), le, 0)     # facing -x, behind: its back
    bs = [S.bsdf(P.EMITTER, (0, 0, 0), emission=le)]
    for target, front in ((5, True), (-5, False)):
        cam = S.camera((0, 0, 0), (np.sign(target), 0, 0), (0, 1, 0), 10.0, 8, 8)
        sc = S.scene([eq, bq], bs, [lt, lb], cam)
        has, hp, hn, hw, dr = O.OracleScene(sc).eye_pass(cam, 1, 1e-4, 0x7A75)
        assert not has.any()
        if front:
            np.testing.assert_allclose(dr, np.tile(f32(le), (64, 1)), rtol=1e-12)
        else:
            assert np.all(dr == 0)


# ---- translucent plane ------------------------------------------------------
def _translucent_plane(a=10.0, h=0.5):
    sa, ss, eta, rho, le = (0.5, 0.3, 0.8), (2.0, 1.5, 2.5), 1.3, (0.9, 0.8, 0.7), (1.0, 2.0, 3.0)
    floor = S.quad((-a, 0, -a), (0, 0, 2 * a), (2 * a, 0, 0))            # normal +y (outward)
    eq, lt = S.emitter_quad((-a, h, -a), (2 * a, 0, 0), (0, 0, 2 * a), le, 1)  # normal -y
    bs = [S.bsdf(P.REFRACTION | P.BSSRDF, rho, sigma_a=sa, sigmap_s=ss, eta=eta),
          S.bsdf(P.EMITTER, (0, 0, 0), emission=le)]
    cam = S.camera((0, h / 2, 0), (0, -1, 0), (0, 0, 1), 10.0, 16, 16)
    return S.scene([(floor, 0), eq], bs, [lt], cam), dict(sa=sa, ss=ss, eta=eta, rho=rho, le=le, a=a, h=h)


def test_subsurface_render_translucent_plane_closed_form():
    sc, p = _translucent_plane()
    a, h = p["a"], p["h"]
    n1 = 400
    g = (np.arange(n1) + 0.5) / n1 * 2 * a - a
    X, Z = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1).astype(np.float32)
    nrm = np.tile(np.float32([0, 1, 0]), (len(pts), 1))
    area = (2 * a) ** 2 / len(pts)
    os_ = O.OracleScene(sc)
    img, st = os_.render(sc.camera, 400000, 8, 0.01 * sc.diagonal(), sss_points=(pts, nrm, area))
    assert st[0] == 0 and st[2] > 0
    # about half of the eye paths are transmitted (reflect probability 0.5)
    assert abs(st[3] / (8 * 256) - 0.5) < 0.05
    got = img.reshape(-1, 3).mean(0)
    # closed form: Schlick at (near) normal incidence for IOR 1.5 gives R0
    R0 = (0.5 / 2.5) ** 2
    # irradiance at the centre of a square Lambertian emitter of radiance Le
    # at height h, as a fraction of the infinite-plane value pi Le
    gg = (np.arange(2000) + 0.5) / 2000 * 2 * a - a
    XX, ZZ = np.meshgrid(gg, gg)
    kappa = (h * h / (h * h + XX ** 2 + ZZ ** 2) ** 2).sum() * (2 * a / 2000) ** 2 / math.pi
    # area integral of the dipole profile by quadrature of the pinned R_d
    r = np.linspace(0.0, 30.0, 30001)
    rd = np.array([O.dipole(f32(p["sa"]), f32(p["ss"]), float(np.float32(p["eta"])), x * x) for x in r])
    I = np.trapezoid(rd * (2 * math.pi * r)[:, None], r, axis=0)
    fdr = O.orc().orc_dipole_fdr(float(np.float32(p["eta"])))
    le = f32(p["le"])
    # stored irradiance E_j is the gather with weight 1/pi: kappa pi Le / pi
    expect = f32(p["rho"]) * R0 * le + (1.0 - R0) * (1.0 - fdr) / math.pi * kappa * le * I
    np.testing.assert_allclose(got / expect, 1.0, atol=0.03)


# ---- the box scene against the independent path tracer ----------------------
def _blocks(img, b=8):
    H, W, _ = img.shape
    return img.reshape(H // b, b, W // b, b, 3).mean((1, 3))


def test_render_agrees_with_independent_path_tracer_on_box():
    s = P.scenegen(P.SCENE_BOX, 1, 0, 64, 64)
    os_ = O.OracleScene(s)
    ppm, _ = os_.render(s.camera, 200000, 32, 0.01 * s.diagonal())
    pt = os_.pathtrace(s.camera, 256)
    # the global mean: noise of both (~0.3 %) plus the darkening a finite
    # radius leaves along the box's edges (~0.3 %)
    assert abs(ppm.mean() / pt.mean() - 1) < 0.015
    b_ppm, b_pt = _blocks(ppm), _blocks(pt)
    rel = (b_ppm - b_pt) / b_pt.mean()
    # 8 x 8 pixel blocks: the photon map's own block noise is ~1.4 % (two seeds)
    assert np.sqrt((rel ** 2).mean()) < 0.025
    assert np.abs(rel).max() < 0.08


# ---- progressive photon mapping with per-point radii (Hachisuka) -------------
def test_hachisuka_render_sphere_furnace_closed_form():
    """The per-point-radius method converges to the same radiance: inside the
    furnace sphere every render point sees L = rho Phi / (pi A (1 - rho))."""
    rho, le = (0.3, 0.5, 0.7), 4.0
    sc, sph = S.sphere_furnace(rho=rho, le=(le,) * 3, res=24)
    A = S.tri_area(sph).sum()
    phi = math.pi * f32(le) * 0.1 * 0.1
    expect = f32(rho) * phi / (math.pi * A * (1.0 - f32(rho)))
    img, st = O.OracleScene(sc).render_hachisuka(sc.camera, 100000, 6)
    assert st[0] > 0
    np.testing.assert_allclose(img.reshape(-1, 3).mean(0) / expect, 1.0, atol=0.03)


def test_hachisuka_and_probabilistic_agree_on_box():
    """Both consistent estimators of the same image: their box-scene renders
    agree within noise and radius bias (and with the path tracer)."""
    s = P.scenegen(P.SCENE_BOX, 1, 0, 32, 32)
    os_ = O.OracleScene(s)
    h, _ = os_.render_hachisuka(s.camera, 100000, 8)
    pt = os_.pathtrace(s.camera, 128)
    assert abs(h.mean() / pt.mean() - 1) < 0.03
