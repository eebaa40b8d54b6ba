"""Progressive photon mapping with per-point radii (Hachisuka) on the GPU
against the oracle's FP64 restatement (sources/progressive_photon_mapping.cc),
through the C ABI.

The oracle walks each render point's candidates in the same order as the
device (deduplicated buckets in z, y, x order, each in photon order), so the
radius and flux sequences agree except where an FP32 photon position lands on
the other side of a radius than the FP64 one: renders agree far inside the
seed-to-seed noise.  The per-pass initial radius (mean extent of the hit
points / mean image side * 8) agrees to FP32 rounding.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.mean(b))


def _gpu(ctx, s, photons, passes, seed=0x7A75, alpha=0.7):
    ctx.upload(s)
    ctx.build_bvh()
    ctx.set_camera(s.camera)
    ctx.begin(photons, passes, 0.01 * s.diagonal(), alpha=alpha, seed=seed, method=P.METHOD_HACHISUKA)
    sts = [ctx.render_pass() for _ in range(passes)]
    return ctx.image(), sts


@pytest.mark.parametrize("scene,sid,N", [(P.SCENE_BOX, 0, 60000), (P.SCENE_TWO_ROOMS, 8, 300000)])
def test_hachisuka_render_matches_oracle(gpu_ctx, scene, sid, N):
    say(f"Hachisuka render vs oracle, scene {scene}")
    s = P.scenegen(scene, 1, sid, 32, 32)
    passes = 4
    g, sts = _gpu(gpu_ctx, s, N, passes)
    os_ = O.OracleScene(s)
    o, _ = os_.render_hachisuka(s.camera, N, passes)
    o2, _ = os_.render_hachisuka(s.camera, N, passes, seed=0x7A76)
    e, noise = rel_rmse(g, o), rel_rmse(o2, o)
    say(f"relative RMSE {e:.3e}, noise {noise:.3e}")
    assert np.all(np.isfinite(g)) and g.min() >= 0.0
    assert e < 0.1 * noise
    assert all(st["radius"] > 0 for st in sts)
    assert sts[-1]["gather_inradius"] > 0


def test_hachisuka_and_probabilistic_agree(gpu_ctx):
    say("Hachisuka vs probabilistic on the box")
    s = P.scenegen(P.SCENE_BOX, 1, 0, 32, 32)
    h, _ = _gpu(gpu_ctx, s, 100000, 8)
    gpu_ctx.begin(100000, 8, 0.01 * s.diagonal())
    for _ in range(8):
        gpu_ctx.render_pass()
    p = gpu_ctx.image()
    assert abs(h.mean() / p.mean() - 1) < 0.03


def test_hachisuka_restarts_cleanly(gpu_ctx):
    say("Hachisuka: a second begin starts from empty render points")
    s = P.scenegen(P.SCENE_BOX, 1, 0, 24, 24)
    a, _ = _gpu(gpu_ctx, s, 20000, 2)
    b, _ = _gpu(gpu_ctx, s, 20000, 2)
    assert np.array_equal(a, b)


def test_hachisuka_rejects_what_it_does_not_take(gpu_ctx):
    s = P.scenegen(P.SCENE_TWO_ROOMS_SSS, 1, 12, 24, 24)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    with pytest.raises(P.PPPMError):
        gpu_ctx.begin(1000, 1, 0.01 * s.diagonal(), method=P.METHOD_HACHISUKA)  # BSSRDF materials
    with pytest.raises(P.PPPMError):
        gpu_ctx.begin(1000, 1, 0.01 * s.diagonal(), method=7)
    s = P.scenegen(P.SCENE_TWO_ROOMS, 1, 8, 24, 24)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.begin(1000, 1, 0.01 * s.diagonal(), method=P.METHOD_HACHISUKA)
    with pytest.raises(P.PPPMError):
        gpu_ctx.proposal_enable(lam=0.5, anchors=4, train_interval=2, train_steps=8, batch=4096,
                                ring_capacity=1 << 20, learning_rate=2e-3, seed=0x5EED)
