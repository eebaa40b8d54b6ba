"""Writes tests/golden/occluded_reference_64.npz: the FP64 oracle image of the
occluded-light scene (two 6^3 rooms joined by a 0.2 x 2 slit, light in room A,
camera and a 200k-triangle displaced sphere in room B) at 64 x 64, the
reference image of the time-to-error study in bench.py.

The oracle's P-KZ render (oracle/oracle.c, FP64) runs 256 passes of 1 M
photons for each of four seeds that no GPU run uses; the file holds their mean
and the standard error of that mean per pixel (from the spread between the
seeds).  About 20 minutes on 8 cores.

Run from the repository root:  python tests/golden/make_occluded_reference.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tatsy-pppm_amd"))
import oracle as O  # noqa: E402
import pppm_amd as P  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "occluded_reference_64.npz")
RES, PHOTONS, PASSES, ALPHA = 64, 1_000_000, 256, 0.7
SEEDS = (0x0AC1E0, 0x0AC1E1, 0x0AC1E2, 0x0AC1E3)


def main():
    s = P.scenegen(P.SCENE_TWO_ROOMS, 1, 100, RES, RES)
    r1 = 0.01 * s.diagonal()
    os_ = O.OracleScene(s)
    imgs = []
    for seed in SEEDS:
        t0 = time.time()
        img, _ = os_.render(s.camera, PHOTONS, PASSES, r1, alpha=ALPHA, ray_eps=1e-4, seed=seed)
        imgs.append(img)
        print(f"seed {seed:#x}: {time.time() - t0:.0f} s, mean {img.mean():.6e}", flush=True)
    imgs = np.stack(imgs)
    mean = imgs.mean(0)
    stderr = imgs.std(0, ddof=1) / np.sqrt(len(SEEDS))
    np.savez_compressed(OUT, image=mean, stderr=stderr,
                        config=np.array([RES, PHOTONS, PASSES, ALPHA, r1], np.float64),
                        seeds=np.array(SEEDS, np.uint64))
    print("wrote", OUT, "relative noise of the mean",
          float(np.sqrt(np.mean(stderr ** 2)) / mean.mean()))


if __name__ == "__main__":
    main()
