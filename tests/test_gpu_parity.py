"""GPU parity: the HIP path through the C ABI against the oracle on the same
seeded inputs, plus size-independent properties at full benchmark sizes.

Bars: bit-exact for the hash grid (integer/index work); stated tolerances for
floating point (FP32 GPU against the FP64 oracle), written next to each check.
"""
import math
import sys

import ctypes as C

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

pytestmark = pytest.mark.gpu

EPS = 1e-4  # continuation-ray offset used throughout (scene units)


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def random_rays(rng, n, lo, hi):
    rays = np.zeros(n, P.RAY_DTYPE)
    rays["origin"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["dir"] = d
    rays["tmin"] = 0.0
    rays["tmax"] = 1e30
    return rays


# ---------------------------------------------------------------------------
# BVH + traversal
# ---------------------------------------------------------------------------
BUILDERS = [P.BVH_LBVH, P.BVH_PLOC]


@pytest.mark.parametrize("builder", BUILDERS)
@pytest.mark.parametrize("kind,freq", [(P.SCENE_BOX, 8), (P.SCENE_BOX_CAUSTIC, 24), (P.SCENE_TWO_ROOMS, 40)])
def test_intersect_matches_brute_force(gpu_ctx, kind, freq, builder):
    say(f"intersect kind={kind} f={freq} builder={builder}")
    s = P.scenegen(kind, 1, freq, 32, 32)
    gpu_ctx.upload(s)
    info = gpu_ctx.build_bvh(builder)
    say(f"bvh {info} tris={s.n_tris}")
    assert info["builder"] == builder
    assert info["nodes"] >= 1 and 3 * info["depth"] + 1 <= 128
    rng = np.random.default_rng(kind * 100 + freq)
    lo, hi = np.array(s.bbox_min) + 0.05, np.array(s.bbox_max) - 0.05
    rays = random_rays(rng, 20000, lo, hi)
    hits = gpu_ctx.intersect(rays)
    os_ = O.OracleScene(s)
    t_ref, tri_ref = os_.intersect(rays.view(np.float32).reshape(-1, 8), brute=True)
    same_tri = hits["tri"] == tri_ref
    # FP32 watertight vs FP64 Moller-Trumbore: ties at shared edges may pick
    # either neighbour; hit/miss must agree and distances must match
    assert np.mean(same_tri) > 0.999, np.mean(same_tri)
    both = (hits["tri"] >= 0) & (tri_ref >= 0)
    assert np.array_equal(hits["tri"] >= 0, tri_ref >= 0)
    np.testing.assert_allclose(hits["t"][both], t_ref[both], rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize("builder", BUILDERS)
def test_watertight_sphere_never_leaks(gpu_ctx, builder):
    say(f"watertight icosphere builder={builder}")
    s = P.scenegen(P.SCENE_BOX, 3, 64, 8, 8)
    keep = s.bsdf_id == s.bsdf_id.max()  # the second sphere (centre 3.5, 1.0, 3.0)
    sub = P.Scene(s.tri[keep], np.zeros(keep.sum(), np.int32), [s.bsdfs[int(s.bsdf_id.max())]], s.lights,
                  s.camera, s.bbox_min, s.bbox_max)
    gpu_ctx.upload(sub)
    gpu_ctx.build_bvh(builder)
    rng = np.random.default_rng(9)
    n = 200000
    rays = np.zeros(n, P.RAY_DTYPE)
    rays["origin"] = np.array([3.5, 1.0, 3.0]) + rng.uniform(-0.3, 0.3, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # include axis-parallel and edge-grazing directions
    d[:6] = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    rays["dir"] = d
    rays["tmax"] = 1e30
    hits = gpu_ctx.intersect(rays)
    assert np.all(hits["tri"] >= 0), int(np.sum(hits["tri"] < 0))


@pytest.mark.parametrize("builder", BUILDERS)
def test_intersect_empty_and_tiny_scenes(gpu_ctx, builder):
    say(f"tiny scenes builder={builder}")
    base = P.scenegen(P.SCENE_BOX, 1, 1, 8, 8)
    for n in (1, 2, 3, 4, 5, 9):
        sub = P.Scene(base.tri[:n], np.zeros(n, np.int32), [base.bsdfs[0]], base.lights, base.camera,
                      base.bbox_min, base.bbox_max)
        gpu_ctx.upload(sub)
        gpu_ctx.build_bvh(builder)
        rng = np.random.default_rng(n)
        rays = random_rays(rng, 4000, np.array([0.1] * 3), np.array([4.9] * 3))
        hits = gpu_ctx.intersect(rays)
        t_ref, tri_ref = O.OracleScene(sub).intersect(rays.view(np.float32).reshape(-1, 8), brute=True)
        assert np.array_equal(hits["tri"] >= 0, tri_ref >= 0)


def _check_bvh(nodes, tris, root, tri_xyz, alone):
    """Every triangle in exactly one leaf, every child box holding what lies
    below it, and the triangles in `alone` (original ids) in leaves of their own."""
    n = len(tri_xyz)
    assert len(tris) == n
    orig = tris[:, 3].copy().view(np.int32)
    assert np.array_equal(np.sort(orig), np.arange(n))  # a permutation of the input
    v = tris.reshape(n, 3, 4)[:, :, :3]
    np.testing.assert_array_equal(v, tri_xyz.reshape(n, 3, 3)[orig])
    lo_t, hi_t = v.min(axis=1), v.max(axis=1)
    alone_pos = np.isin(orig, alone)
    seen = np.zeros(n, np.int32)
    stack = [(root, None)]
    while stack:
        ref, box = stack.pop()
        if ref < 0:
            u = ~ref & 0xFFFFFFFF
            first, cnt = u >> 3, (u & 7) + 1
            seen[first:first + cnt] += 1
            if box is not None:
                assert np.all(lo_t[first:first + cnt] >= box[0]) and np.all(hi_t[first:first + cnt] <= box[1])
            if np.any(alone_pos[first:first + cnt]):
                assert cnt == 1
            continue
        nd = nodes[ref]
        for c in range(4):
            ch = nd.child[c]
            if ch == 0x7FFFFFFF:
                continue
            cb = (np.array([nd.lox[c], nd.loy[c], nd.loz[c]]), np.array([nd.hix[c], nd.hiy[c], nd.hiz[c]]))
            assert np.all(cb[0] <= cb[1])
            if box is not None:
                assert np.all(cb[0] >= box[0]) and np.all(cb[1] <= box[1])
            stack.append((ch, cb))
    assert np.all(seen == 1), (int(np.sum(seen == 0)), int(np.sum(seen > 1)))


@pytest.mark.parametrize("builder", BUILDERS)
@pytest.mark.parametrize("kind,freq", [(P.SCENE_BOX, 8), (P.SCENE_TWO_ROOMS, 24), (P.SCENE_BOX_CAUSTIC, 48)])
def test_bvh_structure(gpu_ctx, kind, freq, builder):
    say(f"bvh structure kind={kind} f={freq} builder={builder}")
    s = P.scenegen(kind, 1, freq, 32, 32)
    gpu_ctx.upload(s)
    info = gpu_ctx.build_bvh(builder)
    nodes, tris, root = gpu_ctx.bvh_download()
    assert info["nodes"] == len(nodes)
    # the room walls: a hundred times the median triangle's area and more,
    # far above the build's big-triangle threshold, so one per leaf
    tri = np.asarray(s.tri, np.float32).reshape(-1, 9)
    e = tri.reshape(-1, 3, 3).max(axis=1) - tri.reshape(-1, 3, 3).min(axis=1)
    area = e[:, 0] * e[:, 1] + e[:, 1] * e[:, 2] + e[:, 2] * e[:, 0]
    alone = np.nonzero(area > 100.0 * np.median(area))[0]
    _check_bvh(nodes, tris, root, tri, alone)


@pytest.mark.parametrize("builder", BUILDERS)
def test_bvh_identical_and_degenerate_triangles(gpu_ctx, builder):
    """Ties everywhere: 2000 copies of one triangle (every pair of clusters at
    the same union area), 200 zero-area triangles and a cluster of tiny ones at
    one point, inside the box's walls; hits and distances against brute force."""
    say(f"identical / degenerate triangles builder={builder}")
    base = P.scenegen(P.SCENE_BOX, 1, 1, 8, 8)
    walls = np.asarray(base.tri, np.float32).reshape(-1, 9)[:12]  # the closed box comes first
    e = walls.reshape(-1, 3, 3).max(axis=1) - walls.reshape(-1, 3, 3).min(axis=1)
    assert np.all(e.max(axis=1) > 4.0)
    one = np.array([[2.0, 1.0, 2.0, 3.0, 1.0, 2.0, 2.5, 2.0, 2.5]], np.float32)
    rng = np.random.default_rng(5)
    # three points on a line, on a 1/8 grid so that the sums are exact: zero
    # area in FP32 and FP64 alike (both tests then miss them)
    line = np.zeros((200, 9), np.float32)
    a = rng.integers(8, 33, (200, 3)).astype(np.float32) / 8
    d = rng.integers(-4, 5, (200, 3)).astype(np.float32) / 8
    line[:, 0:3], line[:, 3:6], line[:, 6:9] = a, a + d, a + 2 * d
    tiny = (np.array([3.5, 3.5, 1.0] * 3, np.float32) + rng.uniform(-1e-4, 1e-4, (500, 9))).astype(np.float32)
    tri = np.concatenate([walls, np.repeat(one, 2000, axis=0), line, tiny]).astype(np.float32)
    sub = P.Scene(tri, np.zeros(len(tri), np.int32), [base.bsdfs[0]], base.lights, base.camera,
                  base.bbox_min, base.bbox_max)
    gpu_ctx.upload(sub)
    info = gpu_ctx.build_bvh(builder)
    assert info["builder"] == builder
    nodes, tris, root = gpu_ctx.bvh_download()
    _check_bvh(nodes, tris, root, tri, np.arange(12))
    rays = random_rays(rng, 20000, np.array([0.1] * 3), np.array([4.9] * 3))
    hits = gpu_ctx.intersect(rays)
    t_ref, tri_ref = O.OracleScene(sub).intersect(rays.view(np.float32).reshape(-1, 8), brute=True)
    # identical triangles tie: which copy is reported is open, the distance is not
    assert np.array_equal(hits["tri"] >= 0, tri_ref >= 0)
    both = (hits["tri"] >= 0) & (tri_ref >= 0)
    np.testing.assert_allclose(hits["t"][both], t_ref[both], rtol=2e-5, atol=2e-6)


def test_ploc_falls_back_on_chain_like_clustering(gpu_ctx):
    """Triangles 1.5 times farther apart each step: every round only the two
    nearest clusters are mutual neighbours, so clustering would take n - 1
    rounds and give a chain; past 64 + 8 log2(n) rounds the radix tree is
    built instead (and reported)."""
    say("PLOC round cap")
    n = 200
    x = (1.5 ** (np.arange(n) - 100.0)).astype(np.float64)
    h = 0.01 * x
    tri = np.zeros((n, 9), np.float32)
    tri[:, 0], tri[:, 3], tri[:, 6] = x, x + h, x
    tri[:, 4], tri[:, 8] = h, h
    base = P.scenegen(P.SCENE_BOX, 1, 1, 8, 8)
    sub = P.Scene(tri, np.zeros(n, np.int32), [base.bsdfs[0]], base.lights, base.camera,
                  base.bbox_min, base.bbox_max)
    gpu_ctx.upload(sub)
    info = gpu_ctx.build_bvh(P.BVH_PLOC)
    assert info["builder"] == P.BVH_LBVH
    nodes, tris, root = gpu_ctx.bvh_download()
    _check_bvh(nodes, tris, root, tri, np.array([], np.int64))


def test_scene_upload_rejects_non_finite_coordinates(gpu_ctx):
    base = P.scenegen(P.SCENE_BOX, 1, 1, 8, 8)
    tri = np.array(base.tri, np.float32).reshape(-1, 9).copy()
    tri[5, 4] = np.nan
    sub = P.Scene(tri, base.bsdf_id, base.bsdfs, base.lights, base.camera, base.bbox_min, base.bbox_max)
    with pytest.raises(P.PPPMError, match="non-finite"):
        gpu_ctx.upload(sub)
    tri[5, 4] = np.inf
    with pytest.raises(P.PPPMError, match="non-finite"):
        gpu_ctx.upload(sub)


# ---------------------------------------------------------------------------
# Hash grid: bit-exact
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["uniform", "boundaries", "wrap", "negative", "duplicates", "single"])
def test_hashgrid_bit_exact_golden(gpu_ctx, golden, name):
    say(f"hashgrid golden {name}")
    pos = golden[f"hg_{name}_pos"]
    prm = golden[f"hg_{name}_params"]
    S, scale, bmin = int(prm[0]), float(prm[1]), prm[2:5]
    g = gpu_ctx.hashgrid(pos, S, scale, bmin)
    o = O.hashgrid_points(pos, S, scale, bmin)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    lists, off = golden[f"hg_{name}_lists"], golden[f"hg_{name}_off"]
    sidx, cs, ce = g
    for i, p in enumerate(pos):
        c = [abs(int((float(p[d]) - bmin[d]) * scale)) for d in range(3)]
        h = O.orc().orc_hash(c[0], c[1], c[2], S)
        assert np.array_equal(sidx[cs[h]:ce[h]], lists[off[i]:off[i + 1]])


@pytest.mark.parametrize("n,S", [(0, 5), (1, 1), (1000, 1), (100000, 99991), (2_000_000, 1_999_993),
                                 (3_000_000, 16_777_259)])
def test_hashgrid_bit_exact_random(gpu_ctx, n, S):
    say(f"hashgrid random n={n} S={S}")
    rng = np.random.default_rng(n + S)
    pos = rng.uniform(0.0, 5.0, (n, 3)).astype(np.float32)
    scale = 1.0 / (2 * 0.0866 * 0.3)
    bmin = np.array([-0.0866, -0.0866, -0.0866])
    g = gpu_ctx.hashgrid(pos, S, scale, bmin)
    o = O.hashgrid_points(pos, S, scale, bmin)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)


def test_hashgrid_boxes_bit_exact(gpu_ctx, golden):
    say("hashgrid boxes")
    bmin_i, bmax_i = golden["hgbox_bmin"], golden["hgbox_bmax"]
    prm = golden["hgbox_params"]
    S, scale, gmin = int(prm[0]), float(prm[1]), prm[2:5]
    g = gpu_ctx.hashgrid_boxes(bmin_i, bmax_i, S, scale, gmin, 10 ** 6)
    o = O.hashgrid_boxes(bmin_i, bmax_i, S, scale, gmin, 10 ** 6)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    with pytest.raises(P.PPPMError) as e:
        gpu_ctx.hashgrid_boxes(bmin_i, bmax_i, S, scale, gmin, 10)
    assert e.value.code == P.ERR_CAPACITY


# ---------------------------------------------------------------------------
# Pass pieces against the oracle (closed box, 1,280-triangle spheres)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c1_small():
    s = P.scenegen(P.SCENE_BOX, 1, 0, 96, 96)
    return s, O.OracleScene(s)


def _render_one_pass(ctx, s, photons, r1, seed=0x7A75, passes=1, **kw):
    ctx.upload(s)
    ctx.build_bvh()
    ctx.set_camera(s.camera)
    ctx.begin(photons, passes, r1, ray_eps=EPS, seed=seed, **kw)
    return [ctx.render_pass() for _ in range(passes)]


def test_photon_pass_matches_oracle(gpu_ctx, c1_small):
    s, os_ = c1_small
    say("photon pass vs oracle")
    N = 30000
    st = _render_one_pass(gpu_ctx, s, N, 0.0866)[0]
    gp, gf, gn, gk = gpu_ctx.last_photons()
    assert st["photons_stored"] == len(gp)
    op, of, on = os_.trace_photons(N, 1, EPS, 0x7A75)
    ok = os_.last_keys
    # keys are sorted (deterministic order) on both sides
    assert np.all(np.diff(gk.astype(np.int64)) > 0)
    common, gi, oi = np.intersect1d(gk, ok, return_indices=True)
    # paths can diverge only where FP32 and FP64 disagree (grazing hits, RR ties)
    assert len(common) > 0.995 * max(len(gk), len(ok)), (len(common), len(gk), len(ok))
    dp = np.abs(gp[gi] - op[oi]).max(axis=1)
    assert np.mean(dp < 1e-3) > 0.995
    rel = np.abs(gf[gi] - of[oi]) / np.abs(of[oi])
    assert np.mean(rel.max(axis=1) < 1e-4) > 0.995
    # oriented normals agree in sign
    assert np.mean(np.sum(gn[gi] * on[oi], axis=1) > 0.999) > 0.995
    # total stored flux within 0.5 %
    np.testing.assert_allclose(gf.sum(axis=0), of.sum(axis=0), rtol=5e-3)


def test_hash_grid_of_a_pass_is_bit_exact(gpu_ctx, c1_small):
    s, _ = c1_small
    say("pass hash grid vs oracle")
    st = _render_one_pass(gpu_ctx, s, 40000, 0.0866)[0]
    gp, _, _, _ = gpu_ctx.last_photons()
    S = st["hash_size"]
    sidx, cs, ce, bmin, scale = gpu_ctx.last_grid(len(gp), S)
    assert scale == 1.0 / (2.0 * 0.0866)
    o = O.hashgrid_points(gp, S, scale, bmin)
    assert np.array_equal(sidx, o[0]) and np.array_equal(cs, o[1]) and np.array_equal(ce, o[2])
    if O.ref_available():
        n = 3000  # the compiled reference HashGrid<int> on the first records
        lists = np.zeros(n * 64, np.int32)
        off = np.zeros(n + 1, np.int64)
        pos = np.ascontiguousarray(gp[:n])
        k = O.ref().ref_hashgrid_point_buckets(O._ptr(pos), n, S, scale, O._ptr(bmin), O._ptr(bmin + 1),
                                               O._ptr(lists), len(lists), O._ptr(off))
        assert k >= 0
        sub = gpu_ctx.hashgrid(pos, S, scale, bmin)
        for i in range(n):
            c = [abs(int((float(pos[i, d]) - bmin[d]) * scale)) for d in range(3)]
            h = O.orc().orc_hash(c[0], c[1], c[2], S)
            assert np.array_equal(sub[0][sub[1][h]:sub[2][h]], lists[off[i]:off[i + 1]])


def test_eye_pass_matches_oracle(gpu_ctx, c1_small):
    s, os_ = c1_small
    say("eye pass vs oracle")
    _render_one_pass(gpu_ctx, s, 1000, 0.0866)
    hp, hn, hw, pix = gpu_ctx.last_hitpoints()
    has, op, on, ow, dr = os_.eye_pass(s.camera, 1, EPS, 0x7A75)
    g_has = np.zeros_like(has)
    g_has[pix] = 1
    assert np.mean(g_has == has) > 0.999
    sel = has[pix] == 1
    p = pix[sel]
    assert np.mean(np.abs(hp[sel] - op[p]).max(axis=1) < 1e-3) > 0.999
    np.testing.assert_allclose(hw[sel], ow[p], rtol=1e-5, atol=1e-7)
    assert np.mean(np.sum(hn[sel] * on[p], axis=1) > 0.9999) > 0.999


@pytest.mark.parametrize("mode", [P.GATHER_WAVE, P.GATHER_TILED, P.GATHER_LANES, P.GATHER_BLOCKS])
def test_gather_matches_oracle(gpu_ctx, c1_small, mode):
    s, os_ = c1_small
    say(f"gather vs oracle mode={mode}")
    pos, flux, nrm = os_.trace_photons(60000, 1, EPS, 0x7A75)
    has, hp, hn, hw, dr = os_.eye_pass(s.camera, 1, EPS, 0x7A75)
    sel = has == 1
    r = 0.0866
    bmin = np.array(s.bbox_min) - r
    S = len(pos)
    g, st = gpu_ctx.gather(pos, flux, nrm, hp[sel], hn[sel], hw[sel], r, S, bmin, mode=mode)
    o, cnt = O.gather(pos, flux, nrm, hp[sel], hn[sel], hw[sel], r, S, bmin)
    assert st["gather_buckets"] == cnt[0] and st["gather_candidates"] == cnt[1]
    # FP32 sums vs FP64; photons on the radius boundary and octahedral normals
    # (16-bit) may flip a few terms of a thousand
    scale = np.abs(o).max(axis=1) + 1e-30
    err = np.abs(g - o).max(axis=1) / scale
    assert np.mean(err < 1e-4) > 0.99, np.quantile(err, [0.5, 0.99, 1.0])
    assert np.all(err < 2e-2)
    assert abs(st["gather_inradius"] - cnt[2]) <= 1e-3 * cnt[2]


# ---------------------------------------------------------------------------
# Whole render against the oracle
# ---------------------------------------------------------------------------
def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.mean(b))


@pytest.mark.parametrize("mode", [P.GATHER_WAVE, P.GATHER_TILED, P.GATHER_LANES, P.GATHER_BLOCKS])
def test_render_matches_oracle_equal_photons(gpu_ctx, c1_small, mode):
    s, os_ = c1_small
    say(f"render vs oracle (4 passes) gather mode={mode}")
    N, passes, r1 = 50000, 4, 0.01 * s.diagonal()
    _render_one_pass(gpu_ctx, s, N, r1, passes=passes, gather_mode=mode)
    g = gpu_ctx.image()
    o, _ = os_.render(s.camera, N, passes, r1, ray_eps=EPS)
    o2, _ = os_.render(s.camera, N, passes, r1, ray_eps=EPS, seed=0x7A76)
    e, noise = rel_rmse(g, o), rel_rmse(o2, o)
    say(f"relative RMSE GPU vs oracle at equal photons: {e:.3e}, seed-to-seed noise {noise:.3e}")
    # same random numbers on both sides: the images differ only where FP32 and
    # FP64 arithmetic take different paths, far inside the Monte Carlo noise
    assert e < 0.05 * noise and e < 1e-2
    assert np.all(np.isfinite(g)) and np.all(g >= 0)


def test_render_is_bitwise_deterministic(gpu_ctx, c1_small):
    s, _ = c1_small
    say("determinism")
    _render_one_pass(gpu_ctx, s, 30000, 0.0866, passes=2)
    a = gpu_ctx.image()
    _render_one_pass(gpu_ctx, s, 30000, 0.0866, passes=2)
    b = gpu_ctx.image()
    assert np.array_equal(a, b)


def test_camera_larger_than_32_bit_pixel_ids_is_rejected(gpu_ctx, c1_small):
    s, _ = c1_small
    say("camera size limit")
    cam = P.Camera()
    C.memmove(C.addressof(cam), C.addressof(s.camera), C.sizeof(P.Camera))
    cam.width, cam.height = 65536, 65536
    with pytest.raises(P.PPPMError) as e:
        gpu_ctx.set_camera(cam)
    assert e.value.code == P.ERR_INVALID
    gpu_ctx.set_camera(s.camera)  # the context stays usable


def test_capacity_error_is_reported(gpu_ctx, c1_small):
    s, _ = c1_small
    say("capacity error")
    with pytest.raises(P.PPPMError) as e:
        _render_one_pass(gpu_ctx, s, 20000, 0.0866, capacity_factor=0.5)
    assert e.value.code == P.ERR_CAPACITY
    # the context stays usable
    _render_one_pass(gpu_ctx, s, 20000, 0.0866)


# ---------------------------------------------------------------------------
# Full size (1024^2, 4M photons, ~0.7M triangles): size-independent properties
# ---------------------------------------------------------------------------
def test_full_size_pass_properties(gpu_ctx):
    say("full size caustic box pass")
    s = P.scenegen(P.SCENE_BOX_CAUSTIC, 1, 0, 1024, 1024)
    r1 = 0.01 * s.diagonal()
    st_l = _render_one_pass(gpu_ctx, s, 4_000_000, r1, gather_mode=P.GATHER_LANES)[0]
    img_lanes = gpu_ctx.image()
    den = np.maximum(np.abs(img_lanes), 1e-6 * img_lanes.max())
    for mode in (P.GATHER_TILED, P.GATHER_WAVE, P.GATHER_BLOCKS):
        # the gather kernels sum the same photon set in different orders
        st = _render_one_pass(gpu_ctx, s, 4_000_000, r1, gather_mode=mode)[0]
        say(f"mode {mode} stats {st}")
        img = gpu_ctx.image()
        assert st_l["gather_inradius"] == st["gather_inradius"]
        assert st_l["gather_candidates"] == st["gather_candidates"]
        if mode == P.GATHER_BLOCKS:
            # per lane the lanes kernel's block decisions
            assert st["gather_tested"] == st_l["gather_tested"] and st["gather_blocks"] == st_l["gather_blocks"]
        else:
            # the whole-bucket kernels test every candidate; the lanes kernel
            # skips the blocks whose box misses the ball, and accepts the same photons
            assert st["gather_tested"] == st["gather_candidates"] and st["gather_blocks"] == 0
        assert 0 < st_l["gather_tested"] < 0.6 * st_l["gather_candidates"] and st_l["gather_blocks"] > 0
        assert np.quantile(np.abs(img - img_lanes) / den, 0.999) < 1e-4
    n = st["photons_stored"]
    assert 1.2 * 4e6 < n < 8 * 4e6
    assert st["hitpoints"] > 0.5 * 1024 * 1024
    img = gpu_ctx.image()
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
    # grid invariants on the real pass: ranges partition 0..n-1 in bucket order,
    # and every record sits in the bucket its FP64 cell hashes to
    gp, _, _, gk = gpu_ctx.last_photons()
    assert np.all(np.diff(gk.astype(np.int64)) > 0)
    S = st["hash_size"]
    sidx, cs, ce, bmin, scale = gpu_ctx.last_grid(n, S)
    assert cs[0] == 0 and ce[-1] == n and np.all(cs[1:] == ce[:-1])
    assert np.array_equal(np.sort(sidx), np.arange(n))
    c = np.abs(np.trunc((gp.astype(np.float64) - bmin) * scale)).astype(np.int64)
    h = ((c[:, 0] * 73856093) ^ (c[:, 1] * 19349663) ^ (c[:, 2] * 83492791)) & 0xFFFFFFFF
    h = h % S
    bucket_of_pos = np.repeat(np.arange(S), ce - cs)
    assert np.array_equal(h[sidx], bucket_of_pos)
    # within each bucket the insertion order is kept (stable)
    same = bucket_of_pos[1:] == bucket_of_pos[:-1]
    assert np.all(sidx[1:][same] > sidx[:-1][same])


@pytest.mark.parametrize("mode", [P.GATHER_LANES, P.GATHER_TILED])
def test_pipelined_passes_equal_single_passes(gpu_ctx, mode):
    # pppm_render_passes overlaps the gather of pass i with the photons of pass
    # i + 1 on three streams; image, photons and counts equal single passes
    say(f"pipelined passes, gather mode {mode}")
    s = P.scenegen(P.SCENE_BOX_CAUSTIC, 3, 10, 96, 64)
    outs = []
    for pipelined in (False, True):
        gpu_ctx.upload(s)
        gpu_ctx.build_bvh()
        gpu_ctx.set_camera(s.camera)
        gpu_ctx.begin(60000, 7, 0.01 * s.diagonal(), ray_eps=EPS, gather_mode=mode)
        if pipelined:
            sts = gpu_ctx.render_passes(2) + gpu_ctx.render_passes(4) + gpu_ctx.render_passes(1)
        else:
            sts = [gpu_ctx.render_pass() for _ in range(7)]
        grid = gpu_ctx.last_grid(sts[-1]["photons_stored"], sts[-1]["hash_size"])
        outs.append((gpu_ctx.image(), gpu_ctx.last_photons(), sts, gpu_ctx.last_hitpoints(), grid))
    assert np.array_equal(outs[0][0], outs[1][0])
    # the last pass's hash grid (record ids in origin order, bucket ranges)
    for a, b in zip(outs[0][4], outs[1][4]):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    for a, b in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(a, b)
    for a, b in zip(outs[0][3], outs[1][3]):
        assert np.array_equal(a, b)
    for x, y in zip(outs[0][2], outs[1][2]):
        for k in ("pass_index", "photons_stored", "photon_segments", "eye_segments", "hitpoints", "gather_candidates",
                  "gather_inradius", "hash_size"):
            assert x[k] == y[k], k
        assert y["gather_ms"] > 0 and y["photon_ms"] > 0


def test_traversal_counting_kernels_change_nothing_but_the_counts(gpu_ctx, c1_small):
    # pppm_count_traversal swaps in kernel variants that count BVH nodes and
    # triangles; images and photon records are those of the plain kernels
    s, _ = c1_small
    say("traversal counting kernels")
    st0 = _render_one_pass(gpu_ctx, s, 40000, 0.0866, passes=2)
    a, pa = gpu_ctx.image(), gpu_ctx.last_photons()
    gpu_ctx.count_traversal(True)
    try:
        st1 = _render_one_pass(gpu_ctx, s, 40000, 0.0866, passes=2)
    finally:
        gpu_ctx.count_traversal(False)
    b, pb = gpu_ctx.image(), gpu_ctx.last_photons()
    assert np.array_equal(a, b)
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)
    for s0, s1 in zip(st0, st1):
        assert s0["photon_nodes"] == 0 and s0["eye_nodes"] == 0
        # every ray visits the root (an inner node for this scene)
        assert s1["photon_nodes"] >= s1["photon_segments"] > 0
        assert s1["eye_nodes"] >= s1["eye_segments"] > 0
        assert 0 < s1["photon_tris"] and 0 < s1["eye_tris"]


# ---------------------------------------------------------------------------
# Image accumulation: FP32 compensated (Kahan) sums against FP64
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["constant", "uniform"])
def test_kahan_accumulation_matches_fp64(gpu_ctx, kind):
    say(f"Kahan accumulator, 10^6 adds ({kind})")
    n = 1_000_000
    if kind == "constant":
        v = np.full(n, 1e-4, np.float32)
    else:
        v = np.random.default_rng(3).uniform(0.0, 2e-4, n).astype(np.float32)
    s, c = gpu_ctx.kahan_probe(v, lanes=64)
    ref = v.astype(np.float64).sum()  # the same FP32 addends summed in FP64
    got = s.astype(np.float64) - c.astype(np.float64)
    # compensated: within a few FP32 ulps of the total
    assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.float32(ref))), (got - ref).max()
    # the plain FP32 running sum of a fifth of the addends already misses that
    # bound (the test can tell compensated from plain accumulation)
    naive = np.float32(0.0)
    for x in v[:200_000]:
        naive = np.float32(naive + x)
    assert abs(float(naive) - v[:200_000].astype(np.float64).sum()) > 4 * np.spacing(np.float32(ref))
