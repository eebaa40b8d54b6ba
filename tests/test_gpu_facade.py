"""The reference-shaped C++ interface (tatsy-pppm_amd/host/pppm_amd.h) on the
GPU: Scene / QBVHAccel / HashGrid / ProgressivePhotonMappingProb driven the
way the reference's application drives them, and its image identical to the
C ABI path with the same settings."""
import os
import subprocess

import numpy as np
import pytest

import pppm_amd as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_main")


def test_cpp_interface_matches_c_abi(gpu_ctx, tmp_path):
    print("[case] C++ interface", flush=True)
    out = tmp_path / "facade.bin"
    r = subprocess.run([BIN, str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "OK" in r.stdout
    img = np.fromfile(out, np.float64).reshape(64, 64, 3)
    s = P.scenegen(P.SCENE_BOX, 1, 0, 64, 64)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.begin(20000, 2, 0.01 * s.diagonal(), ray_eps=1e-4, seed=0x7A75)
    gpu_ctx.render_pass()
    gpu_ctx.render_pass()
    ref = gpu_ctx.image()
    assert np.array_equal(img, ref)


def test_cpp_translucent_and_learned_renders_match_c_abi(gpu_ctx, tmp_path):
    print("[case] C++ interface: BSSRDF and learned proposal", flush=True)
    out = [tmp_path / n for n in ("a.bin", "sss.bin", "prop.bin", "ppm.bin")]
    r = subprocess.run([BIN] + [str(o) for o in out], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "OK" in r.stdout
    s = P.scenegen(P.SCENE_TWO_ROOMS_SSS, 1, 12, 48, 48)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.subsurface_configure(20000)
    gpu_ctx.begin(100000, 2, 0.01 * s.diagonal(), ray_eps=1e-4, seed=0x7A75)
    for _ in range(2):
        gpu_ctx.render_pass()
    assert np.array_equal(np.fromfile(out[1], np.float64).reshape(48, 48, 3), gpu_ctx.image())
    gpu_ctx.subsurface_configure(0)
    s = P.scenegen(P.SCENE_TWO_ROOMS, 1, 8, 48, 48)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.begin(50000, 4, 0.01 * s.diagonal(), ray_eps=1e-4, seed=0x7A75)
    gpu_ctx.proposal_enable(lam=0.5, anchors=4, train_interval=2, train_steps=8, batch=4096, ring_capacity=1 << 20,
                            learning_rate=2e-3, seed=0x5EED)
    for _ in range(4):
        gpu_ctx.render_pass()
    assert np.array_equal(np.fromfile(out[2], np.float64).reshape(48, 48, 3), gpu_ctx.image())
    # ProgressivePhotonMapping: the C ABI with method = PPPM_METHOD_HACHISUKA
    s = P.scenegen(P.SCENE_BOX, 1, 0, 40, 32)
    gpu_ctx.upload(s)
    gpu_ctx.build_bvh()
    gpu_ctx.set_camera(s.camera)
    gpu_ctx.begin(30000, 3, 0.01 * s.diagonal(), ray_eps=1e-4, seed=0x7A75, method=P.METHOD_HACHISUKA)
    for _ in range(3):
        gpu_ctx.render_pass()
    assert np.array_equal(np.fromfile(out[3], np.float64).reshape(32, 40, 3), gpu_ctx.image())


def test_reference_application_shape_runs(tmp_path):
    """sources/main/appmain.cc:11-26's call pattern (default Camera assigned in
    setScene, a placed .ply mesh, render with the default sampler) on the GPU."""
    print("[case] application-shaped program", flush=True)
    exe = os.path.join(ROOT, "tests", "cpp", "appmain_shape")
    r = subprocess.run([exe, "64", "48", "2"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("OK")
