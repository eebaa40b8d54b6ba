"""GPU parity of the batched, graph-captured pass loop through the C ABI.

Bar: every batched scene's image equals the image of the same scene rendered
alone with pppm_render_pass (same keys, hash size and gather kernel) bit for
bit; solo renders are the oracle here.
"""
import sys

import numpy as np
import pytest

import pppm_amd as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


SCENES = [  # kind, seed, frequency, resolution, photons
    (P.SCENE_BOX, 1, 4, 48, 30000),
    (P.SCENE_BOX_CAUSTIC, 2, 8, 64, 60000),
    (P.SCENE_RANDOM, 3, 12, 32, 20000),
    (P.SCENE_TWO_ROOMS, 4, 6, 40, 50000),
    (P.SCENE_BOX, 5, 2, 16, 5000),
]


def _solo(ctx, s, photons, passes, r1, hash_size, mode, seed):
    ctx.upload(s)
    ctx.build_bvh()
    ctx.set_camera(s.camera)
    ctx.begin(photons, passes, r1, seed=seed, hash_size=hash_size, gather_mode=mode, capacity_factor=4.0)
    for _ in range(passes):
        ctx.render_pass()
    return ctx.image()


@pytest.mark.parametrize("tune", [False, True])
def test_batched_scenes_equal_solo_renders(gpu_ctx, tune):
    say(f"batch vs solo, tune={tune}")
    passes = 4
    b = P.Batch(branches=3)
    scenes = []
    for k, (kind, seed, f, res, n) in enumerate(SCENES):
        s = P.scenegen(kind, seed, f, res, res)
        r1 = 0.01 * s.diagonal()
        b.add(s, n, passes, r1, hash_size=n, seed=0x100 + k, capacity_factor=4.0)
        scenes.append((s, n, r1, 0x100 + k))
    if tune:
        b.tune()
    b.run(2)
    b.run(passes - 2)  # the captured graph is replayed across calls
    for k, (s, n, r1, seed) in enumerate(scenes):
        info = b.info(k)
        got = b.image(k)
        want = _solo(gpu_ctx, s, n, passes, r1, n, info["gather_mode"], seed)
        assert np.array_equal(got, want), (k, info, float(np.abs(got - want).max()))
        # the tuned traversal shape changes the schedule, never the image
        assert info["trace_keep"] in (8, 16, 24, 32, 48)
        assert got.max() > 0.0
    b.close()


def test_batch_reports_capacity_overflow():
    say("batch capacity overflow")
    b = P.Batch(branches=1)
    s = P.scenegen(P.SCENE_BOX, 1, 2, 16, 16)
    b.add(s, 20000, 2, 0.05, hash_size=20000, capacity_factor=0.25)
    with pytest.raises(P.PPPMError) as e:
        b.run(1)
    assert e.value.code == P.ERR_CAPACITY
    b.close()


def test_batch_rejects_bad_input():
    b = P.Batch(branches=2)
    s = P.scenegen(P.SCENE_BOX, 1, 2, 16, 16)
    with pytest.raises(P.PPPMError):
        b.add(s, 1000, 2, 0.05, hash_size=0)
    with pytest.raises(P.PPPMError):
        b.run(1)  # no scenes
    b.add(s, 1000, 2, 0.05, hash_size=1000)
    with pytest.raises(P.PPPMError):
        b.run(3)  # more passes than begun with
    b.close()


def test_batched_scene_equals_solo_at_power_of_two_hash_size(gpu_ctx):
    """A hash size of exactly 2^23: single and enqueued passes must order each
    bucket's records alike (the in-bucket Morton bits are derived the same way)."""
    say("batch vs solo at hash size 2^23")
    passes, S = 2, 1 << 23
    s = P.scenegen(P.SCENE_BOX, 3, 6, 48, 48)
    n, r1 = 40000, 0.01 * s.diagonal()
    b = P.Batch(branches=1)
    b.add(s, n, passes, r1, hash_size=S, seed=0x321, capacity_factor=4.0, gather_mode=P.GATHER_LANES)
    b.run(passes)
    got = b.image(0)
    want = _solo(gpu_ctx, s, n, passes, r1, S, P.GATHER_LANES, 0x321)
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    b.close()
