"""GPU parity of the subsurface path (dipole BSSRDF) through the C ABI.

Bars: the dipole sum within 2e-5 relative (FP32 with Kahan sums against the
FP64 oracle's brute force); the uniform-plane quadrature within 2e-3 of the
closed-form area integral; surface irradiance within 1e-4 relative for
99.5 % of the points against the oracle's gather over the same photons.
"""
import math
import sys

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

pytestmark = pytest.mark.gpu

MARBLE = ((0.0021, 0.0041, 0.0071), (2.19, 2.62, 3.00))


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _material(sa, ss, eta=1.3):
    b = P.Bsdf()
    b.type = P.REFRACTION | P.BSSRDF
    for k in range(3):
        b.reflectance[k] = 0.99
        b.sigma_a[k] = sa[k]
        b.sigmap_s[k] = ss[k]
    b.ior = 1.5
    b.eta = eta
    return b


@pytest.mark.parametrize("n_pts,n_q", [(3000, 700), (1, 1), (257, 255)])
def test_dipole_sum_matches_oracle(gpu_ctx, n_pts, n_q):
    say(f"dipole sum {n_pts} x {n_q}")
    rng = np.random.default_rng(n_pts)
    sa, ss = (0.01, 0.02, 0.04), (1.0, 1.5, 0.7)
    pts = rng.uniform(-1.5, 1.5, (n_pts, 3)).astype(np.float32)
    E = rng.random((n_pts, 3)).astype(np.float32)
    q = rng.uniform(-1.5, 1.5, (n_q, 3)).astype(np.float32)
    g = gpu_ctx.dipole_eval(_material(sa, ss), pts, E, 0.003, q)
    o = O.dipole_sum(sa, ss, 1.3, pts, E, 0.003, q)
    np.testing.assert_allclose(g, o, rtol=2e-5, atol=1e-12)


def test_uniform_plane_matches_closed_form(gpu_ctx):
    say("dipole uniform plane")
    h = 0.01
    xs = (np.arange(-250, 250) + 0.5) * h         # a 5 x 5 plane, beyond r* of marble
    X, Z = np.meshgrid(xs, xs)
    pts = np.column_stack([X.ravel(), np.zeros(X.size), Z.ravel()]).astype(np.float32)
    E = np.ones_like(pts)
    L = gpu_ctx.dipole_eval(_material(*MARBLE), pts, E, h * h, np.zeros((1, 3), np.float32))[0]
    F = O.orc().orc_dipole_fdr(1.3)
    A = (1 + F) / (1 - F)
    for c in range(3):
        sa, ss = MARBLE[0][c], MARBLE[1][c]
        st = sa + ss
        s = math.sqrt(3 * sa * st)
        ap, zp = ss / st, 1 / st
        zn = zp * (1 + 4 / 3 * A)
        lo, hi = 0.0, 100.0
        for _ in range(200):
            m = 0.5 * (lo + hi)
            if O.dipole(MARBLE[0], MARBLE[1], 1.3, m * m)[c] > 0:
                lo = m
            else:
                hi = m
        dp, dn = math.hypot(lo, zp), math.hypot(lo, zn)
        integral = ap / (2 * s) * (math.exp(-s * zp) - zp * math.exp(-s * dp) / dp - math.exp(-s * zn)
                                   + zn * math.exp(-s * dn) / dn)
        assert L[c] == pytest.approx((1 - F) / math.pi * integral, rel=2e-3)


@pytest.fixture(scope="module")
def sss_scene():
    return P.scenegen(P.SCENE_TWO_ROOMS_SSS, 1, 24, 64, 64)


def _begin(ctx, s, photons, passes, points=20000, seed=0x7A75):
    ctx.upload(s)
    ctx.build_bvh()
    ctx.set_camera(s.camera)
    ctx.subsurface_configure(points)
    ctx.begin(photons, passes, 0.01 * s.diagonal(), seed=seed)


def test_surface_points_cover_the_translucent_mesh(gpu_ctx, sss_scene):
    s = sss_scene
    say("surface points")
    _begin(gpu_ctx, s, 1000, 1)
    pos, nrm, area = gpu_ctx.subsurface_points()
    assert len(pos) == 20000
    sm = int(s.bsdf_id[-1])
    tri = s.tri[s.bsdf_id == sm].reshape(-1, 3, 3).astype(np.float64)
    tot = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    assert area * len(pos) == pytest.approx(tot, rel=1e-6)
    c = np.array([9.0, 1.3, 3.0])
    r = np.linalg.norm(pos - c, axis=1)
    assert r.min() > 1.3 * 0.9 and r.max() < 1.3 * 1.1      # on the displaced sphere
    assert np.mean(np.sum(nrm * (pos - c), axis=1) > 0) > 0.99  # outward normals


def test_irradiance_matches_oracle_gather(gpu_ctx, sss_scene):
    s = sss_scene
    say("surface irradiance")
    _begin(gpu_ctx, s, 2000000, 1)
    st = gpu_ctx.render_pass()
    assert st["irradiance_photons"] > 500 and st["sss_points"] == 20000
    E = gpu_ctx.subsurface_irradiance()
    ip, iflux, inrm, _ = gpu_ctx.last_irradiance_photons()
    pos, nrm, _ = gpu_ctx.subsurface_points()
    w = np.full_like(pos, 1.0 / math.pi)
    S = st["irradiance_photons"]
    bmin = np.array(s.bbox_min) - 0.01 * s.diagonal()
    o, _ = O.gather(ip, iflux, inrm, pos, nrm, w, st["radius"], S, bmin)
    ok = np.all(np.abs(E - o) <= 1e-4 * np.abs(o) + 1e-12, axis=1)
    assert ok.mean() > 0.995, ok.mean()
    assert E.max() > 0


def test_sss_render_is_finite_and_deterministic(gpu_ctx, sss_scene):
    s = sss_scene
    say("sss render")
    imgs = []
    for _ in range(2):
        _begin(gpu_ctx, s, 200000, 2)
        sts = [gpu_ctx.render_pass() for _ in range(2)]
        imgs.append(gpu_ctx.image())
    assert all(t["sss_hits"] > 0 for t in sts)
    assert np.array_equal(imgs[0], imgs[1])
    assert np.all(np.isfinite(imgs[0])) and imgs[0].min() >= 0.0


# ---- the octree cut (pppm_subsurface_cut) ------------------------------------
@pytest.mark.parametrize("name", ["sphere", "plane", "sphere_tight"])
def test_dipole_cut_matches_oracle_octree(gpu_ctx, golden, name):
    """The GPU's walk of the cut against the oracle's FP64 octree (itself
    pinned to the reference's Octree) on the same points and queries; the
    FP32 error test can take or open a node the FP64 one does not, which moves
    a query by a fraction of that node's share."""
    say(f"dipole octree cut, {name}")
    pts, E, q = golden[f"oct_{name}_pts"], golden[f"oct_{name}_E"], golden[f"oct_{name}_q"]
    prm = golden[f"oct_{name}_params"]
    sa, ss, eta, area, eps = prm[0:3], prm[3:6], prm[6], prm[7], prm[8]
    try:
        gpu_ctx.subsurface_cut(eps)
        g = gpu_ctx.dipole_eval(_material(sa, ss, eta), pts, E, area, q)
    finally:
        gpu_ctx.subsurface_cut(0.0)
    o, _ = O.dipole_octree(sa, ss, eta, pts, E, area, q, max_error=eps)
    big = o.sum(1) > 1e-3 * o.sum(1).max()
    rel = np.abs(g[big] / o[big] - 1)
    say(f"median {np.median(rel):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 2e-5 and rel.max() < 2e-3


def test_dipole_cut_without_error_is_the_exact_sum(gpu_ctx):
    say("dipole cut at max_error 1e-30")
    rng = np.random.default_rng(4)
    sa, ss = (0.01, 0.02, 0.04), (1.0, 1.5, 0.7)
    pts = rng.uniform(-1.5, 1.5, (3000, 3)).astype(np.float32)
    pts[100:110] = pts[5]  # coincident points: one leaf
    E = rng.random((3000, 3)).astype(np.float32)
    q = rng.uniform(-1.5, 1.5, (500, 3)).astype(np.float32)
    try:
        gpu_ctx.subsurface_cut(1e-30)
        g = gpu_ctx.dipole_eval(_material(sa, ss), pts, E, 0.003, q)
    finally:
        gpu_ctx.subsurface_cut(0.0)
    o = O.dipole_sum(sa, ss, 1.3, pts, E, 0.003, q)
    np.testing.assert_allclose(g, o, rtol=2e-5, atol=1e-12)
    with pytest.raises(P.PPPMError):
        gpu_ctx.subsurface_cut(-1.0)


def test_sss_render_with_cut_is_close_to_the_exact_sum(gpu_ctx, sss_scene):
    """The whole render with the reference's cut (0.05) against the exact sum:
    same photons and eye paths, so the images differ only by the cut."""
    s = sss_scene
    say("sss render, cut vs exact")
    imgs = []
    for eps in (0.0, 0.05):
        gpu_ctx.subsurface_cut(eps)
        _begin(gpu_ctx, s, 200000, 2)
        sts = [gpu_ctx.render_pass() for _ in range(2)]
        imgs.append(gpu_ctx.image())
    gpu_ctx.subsurface_cut(0.0)
    assert sts[-1]["sss_hits"] > 0
    exact, cut = imgs
    d = np.abs(cut - exact).sum() / exact.sum()
    say(f"relative L1 difference {d:.2e}")
    assert np.all(np.isfinite(cut)) and cut.min() >= 0.0
    assert d < 5e-3
