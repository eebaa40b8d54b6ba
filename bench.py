"""Benchmark: progressive photon mapping passes on one MI355X.

Workload (one GPU): the caustic box -- a closed 5^3 box, two value-noise
displaced icospheres (327,680 triangles each), a mirror and an IOR-1.5 glass
sphere (20,480 each), one 1x1 area light; 1024x1024 pinhole camera,
4,000,000 photons per pass, alpha = 0.7, r_1 = 1 % of the scene diagonal.
A step is one full pass: photon tracing, deterministic photon ordering, hash
grid build, eye pass, gather and Kahan accumulation into the image.

Prints one JSON line.  value = emitted photons per second of whole-pass wall
time with the scene resident in HBM.  The roofline object covers the gather
kernel, timed with HIP events on the library's stream; cpu_baseline times the
oracle (FP64 CPU restatement, OpenMP) on one full pass of the same workload.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tatsy-pppm_amd"))

METRIC = "time to RMSE 1e-2 vs FP64 oracle (plain vs learned); photons/s; gather GB/s"
HBM_PEAK_GBPS = 8000.0
HIT_RECORD_BYTES = 40       # position 12 + normal 12 + weight 12 + pixel 4
ACCUM_BYTES = 48            # FP32 sum + compensation, read and written
BUCKET_RANGE_BYTES = 8      # cell start + end
PHOTON_BYTES = 24           # position 12 + flux 12


def gather_bytes(st):
    """Algorithmic bytes of one gather launch (reference cell convention)."""
    return (st["hitpoints"] * (HIT_RECORD_BYTES + ACCUM_BYTES)
            + BUCKET_RANGE_BYTES * st["gather_buckets"] + PHOTON_BYTES * st["gather_candidates"])


def read_traffic():
    """HBM bytes per gather launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "gather_pmc_caustic_box_1024.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(scene, photons, r1, ray_eps, seed):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    t0 = time.perf_counter()
    os_ = O.OracleScene(scene)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    img, st = os_.render(scene.camera, photons, 1, r1, ray_eps=ray_eps, seed=seed)
    t_pass = time.perf_counter() - t0
    return {
        "value": photons / t_pass,
        "unit": "photons/s",
        "cores": int(O.orc().orc_num_threads()),
        "kind": "port",
        "sample": (f"one full pass of the same workload on the host CPU ({cpu_model()}): "
                   f"{photons} photons, {scene.camera.width}x{scene.camera.height} eye paths, "
                   f"FP64 OpenMP oracle; pass {t_pass:.2f} s, QBVH build {t_build:.2f} s excluded"),
    }, img


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=4_000_000)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--gather-mode", choices=["auto", "lanes", "tiled", "wave"], default="auto")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="after the passes, copy BYTES device-to-device once (a plain copy kernel "
                         "of known traffic that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}). "
              "Passes shard across GPUs as independent photon streams; see the design notes.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pppm_amd as P

    scene = P.scenegen(P.SCENE_BOX_CAUSTIC, args.seed, 0, args.width, args.height)
    r1 = 0.01 * scene.diagonal()
    ray_eps = 1e-4
    seed = 0x7A75
    ctx = P.Context(0)
    ctx.upload(scene)
    bvh = ctx.build_bvh()
    ctx.set_camera(scene.camera)
    total = args.warmup + args.steps
    mode = {"auto": P.GATHER_AUTO, "lanes": P.GATHER_LANES, "tiled": P.GATHER_TILED,
            "wave": P.GATHER_WAVE}[args.gather_mode]
    ctx.begin(args.photons, total, r1, ray_eps=ray_eps, seed=seed, gather_mode=mode)
    for _ in range(args.warmup):
        ctx.render_pass()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(ctx.render_pass())  # synchronous: returns after the pass completes
    elapsed = time.perf_counter() - t0
    img = ctx.image()
    assert np.all(np.isfinite(img))

    photons_per_s = args.photons * args.steps / elapsed
    segs = sum(s["photon_segments"] + s["eye_segments"] for s in stats)
    trace_ms = sum(s["photon_ms"] + s["eye_ms"] for s in stats)
    gather_ms = sum(s["gather_ms"] for s in stats) / args.steps
    gbytes = sum(gather_bytes(s) for s in stats) / args.steps
    achieved = gbytes / (gather_ms * 1e-3) / 1e9
    photon_kernel_ms = sum(s["photon_ms"] for s in stats) / args.steps
    out = {
        "metric": METRIC,
        "value": photons_per_s,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "caustic box: 2 displaced icospheres f=128 + mirror + glass, 1 area light",
            "triangles": int(scene.n_tris),
            "resolution": [args.width, args.height],
            "photons_per_pass": args.photons,
            "alpha": 0.7,
            "radius1": r1,
            "bvh": {k: bvh[k] for k in ("nodes", "leaves", "depth", "build_ms")},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_gather",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": read_traffic(),
            "algorithmic_bytes_per_launch": gbytes,
            "avg_launch_ms": gather_ms,
        },
        "rays_per_s": segs / (trace_ms * 1e-3),
        "photon_kernel_photons_per_s": args.photons / (photon_kernel_ms * 1e-3),
        "gather_gbps": achieved,
        "stage_ms": {k: sum(s[k] for s in stats) / args.steps
                     for k in ("photon_ms", "order_ms", "grid_ms", "eye_ms", "hitsort_ms", "gather_ms",
                               "total_ms")},
        "gather_mode": args.gather_mode,
        "photons_stored_per_pass": sum(s["photons_stored"] for s in stats) / args.steps,
        "hitpoints_per_pass": sum(s["hitpoints"] for s in stats) / args.steps,
        "time_to_rmse": None,
    }
    if args.copy_calibration > 0:
        copy_ms = ctx.stream_copy(args.copy_calibration)
        out["copy_calibration"] = {"bytes": args.copy_calibration, "ms": copy_ms,
                                   "GBps": 2 * args.copy_calibration / (copy_ms * 1e-3) / 1e9}
    ctx.close()
    if not args.no_cpu:
        cb, _ = cpu_baseline(scene, args.photons, r1, ray_eps, seed)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
