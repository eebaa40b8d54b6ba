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
