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
whole_job renders all 256 passes of the workload in one pipelined call and
reports its photons/s and the gather's useful-byte fraction by pass range.
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



# Proposal retraining: Adam steps per retrain and samples per step (set from
# the command line)
TRAIN = [8, 65536]

def gather_bytes(st):
    """Algorithmic bytes of one gather launch (reference cell convention)."""
    return (st["hitpoints"] * (HIT_RECORD_BYTES + ACCUM_BYTES)
            + BUCKET_RANGE_BYTES * st["gather_buckets"] + PHOTON_BYTES * st["gather_candidates"])


def useful_bytes(st):
    """Gather bytes if only in-radius photons were read."""
    return st["hitpoints"] * (HIT_RECORD_BYTES + ACCUM_BYTES) + PHOTON_BYTES * st["gather_inradius"]


def traversal_bytes(rays, nodes_per_ray, tris_per_ray):
    """Algorithmic bytes of one photon-pass launch: ray in / hit out, nodes, triangles."""
    return rays * (32 + 16 + 128 * nodes_per_ray + 36 * tris_per_ray)


def read_traffic(name="gather_pmc_caustic_box_1024.txt"):
    """HBM bytes per launch from a committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def read_sq(kernel, counter, name="sq_caustic_box_1024.txt"):
    """Per-launch value of an SQ counter for one kernel from a committed
    rocprofv3 PMC summary (None when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    current = None
    for line in open(path):
        if line.startswith("kernel "):
            current = line[len("kernel "):].split(":", 1)[0]
        elif current == kernel and line.strip().startswith(counter + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def issue_bound(kernel, probe, cus, iso_ms, pipelined_ms):
    """The kernel against the VALU issue roofline: its VALU wave instructions
    per launch (PMC summary) at the measured cycles per full-rate wave
    instruction, over four SIMDs per CU at the reported clock, against its
    launch time alone and in the pipeline.  lane_utilization = active lanes per
    VALU instruction / 64 (SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64))."""
    valu = read_sq(kernel, "SQ_INSTS_VALU_per_launch")
    if valu is None or not iso_ms:
        return None
    issue_ms = 1e3 * valu * probe["valu_cycles"] / (4.0 * cus * probe["clock_hz"])
    return {
        "bound": "valu",
        "kernel": kernel,
        "valu_instr_per_launch": valu,
        "cycles_per_instr": probe["valu_cycles"],
        "issue_ms": issue_ms,
        "frac_isolated": issue_ms / iso_ms,
        "frac_pipelined": issue_ms / pipelined_ms if pipelined_ms else None,
        "lane_utilization": read_sq(kernel, "lane_utilization"),
        "source": "profiles/sq_caustic_box_1024.txt (VALU count, lane utilisation); pppm_issue_probe (cycles, clock)",
    }


def cpu_model():
    """Model name, family / model / stepping, and the logical CPUs of the host."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
    except OSError:
        pass
    name = info.get("model name", platform.processor())
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}), {os.cpu_count()} logical CPUs")


def smt_state():
    """SMT as the kernel reports it (on / off / notsupported), with threads per core."""
    state = "unknown"
    try:
        state = open("/sys/devices/system/cpu/smt/control").read().strip()
    except OSError:
        pass
    try:
        sib = open("/sys/devices/system/cpu/cpu0/topology/thread_siblings_list").read().strip()
        n = sum(len(range(int(a), int(b) + 1)) if "-" in r else 1
                for r in sib.split(",") for a, _, b in [r.partition("-")])
        state += f", {n} threads per core"
    except (OSError, ValueError):
        pass
    return state


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
        "cpu": cpu_model(),
        "smt": smt_state(),
        "sample": (f"one full pass of the same workload on the host CPU ({cpu_model()}, SMT {smt_state()}): "
                   f"{photons} photons, {scene.camera.width}x{scene.camera.height} eye paths, "
                   f"FP64 OpenMP oracle; pass {t_pass:.2f} s, QBVH build {t_build:.2f} s excluded"),
    }, img


def error_study(P, np, photons, width, passes, ref_passes, lam, probe_every=8):
    """Time to equal error, plain against learned emission, on the occluded-light
    scene (two 6^3 rooms joined by a 0.2 x 2 door slit, light in room A, camera
    and a ~200k-triangle displaced sphere in room B).

    Error = RMSE against the mean of two independent long plain runs, with the
    references' own noise removed (their half-difference estimates it), relative
    to the reference mean.  The threshold is plain's error after `passes`
    passes; the learned run's time to reach it is interpolated between probes.
    The noise part alone (no bias) comes from two seeds of each method at
    `passes` passes: variance_ratio = plain / learned, and noise_only_speedup
    scales it by the per-pass time ratio.
    """
    scene = P.scenegen(P.SCENE_TWO_ROOMS, 1, 100, width, width)
    r1 = 0.01 * scene.diagonal()
    ctx = P.Context(0)
    ctx.upload(scene)
    ctx.build_bvh()
    ctx.set_camera(scene.camera)

    def run(seed, n, learned, ref=None):
        ctx.begin(photons, n, r1, seed=seed)
        if learned:
            ctx.proposal_enable(lam=lam, anchors=8, train_interval=8, train_steps=TRAIN[0], batch=TRAIN[1],
                                ring_capacity=16 << 20, learning_rate=2e-3, seed=0x5EED)
        t, curve, extra = 0.0, [], 0.0
        for i in range(1, n + 1):
            t0 = time.perf_counter()
            st = ctx.render_pass()
            t += time.perf_counter() - t0
            extra += st["proposal_ms"]
            if ref is not None and (i % probe_every == 0 or i == 1):
                d = ctx.image() - ref
                curve.append((i, t, float(np.mean(d * d))))
        return ctx.image(), t, curve, extra / n

    ref1, t_ref, _, _ = run(0xBEE1, ref_passes, False)
    ref2, _, _, _ = run(0xBEE2, ref_passes, False)
    ref = 0.5 * (ref1 + ref2)
    ref_noise = float(np.mean((ref1 - ref2) ** 2)) / 4.0
    mean_ref = float(ref.mean())

    def rel(mse):
        return float(np.sqrt(max(mse - ref_noise, 0.0))) / mean_ref

    img_pa, t_plain, c_plain, _ = run(0xA0A, passes, False, ref)
    img_la, t_learned, c_learned, prop_ms = run(0xA0A, passes, True, ref)
    # noise alone (no bias): half the mean squared difference of two seeds
    img_pb, _, _, _ = run(0xB0B, passes, False)
    img_lb, _, _, _ = run(0xB0B, passes, True)
    noise_plain = float(np.mean((img_pa - img_pb) ** 2)) / 2.0
    noise_learned = float(np.mean((img_la - img_lb) ** 2)) / 2.0
    target = rel(c_plain[-1][2])
    reach = None
    prev = None
    for i, t, m in c_learned:
        e = rel(m)
        if e <= target:
            if prev is None:
                reach = t
            else:
                (t0, e0) = prev
                reach = t0 + (t - t0) * (e0 - target) / max(e0 - e, 1e-30)
            break
        prev = (t, e)
    learned_at_plain_time = None
    for (i, t, m) in c_learned:
        if t >= t_plain:
            learned_at_plain_time = rel(m)
            break
    ctx.close()
    return {
        "scene": "two 6^3 rooms, 0.2 x 2 door slit, light in room A, camera + displaced sphere in room B",
        "triangles": int(scene.n_tris),
        "resolution": [width, width],
        "photons_per_pass": photons,
        "error": "RMSE vs reference / mean(reference), reference noise removed",
        "reference": f"mean of 2 independent plain GPU runs x {ref_passes} passes (noise estimate subtracted)",
        "reference_rel_noise": float(np.sqrt(ref_noise)) / mean_ref,
        "reference_mean_radiance": mean_ref,
        "abs_rmse_plain_pass1": float(np.sqrt(c_plain[0][2])),
        "threshold_rel": target,
        "plain_s": t_plain,
        "plain_passes": passes,
        "learned_s": reach,
        "speedup": (t_plain / reach) if reach else None,
        "learned_rel_at_plain_time": learned_at_plain_time,
        "learned_total_s": t_learned,
        "learned_overhead_ms_per_pass": prop_ms,
        "lambda": lam,
        "noise_rel_plain": float(np.sqrt(noise_plain)) / mean_ref,
        "noise_rel_learned": float(np.sqrt(noise_learned)) / mean_ref,
        "variance_ratio": noise_plain / max(noise_learned, 1e-300),
        "noise_only_speedup": (noise_plain / max(noise_learned, 1e-300)) * (t_plain / t_learned),
        "curve_plain": [(i, round(t, 4), round(rel(m), 5)) for i, t, m in c_plain],
        "curve_learned": [(i, round(t, 4), round(rel(m), 5)) for i, t, m in c_learned],
    }


def c1_parity_run(P, np):
    """The reference's CPU-runnable case (C1: closed box, two f = 8 spheres,
    256^2, 200k photons per pass, 32 passes, alpha 0.7, r1 = 1 % of the
    diagonal) in full, on the GPU and with the FP64 oracle on the host cores,
    same seed: the CPU path timed beside the GPU, and the relative RMSE between
    the two images at an equal photon count (both use the same random numbers,
    so it measures only FP32 against FP64 arithmetic)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    scene = P.scenegen(P.SCENE_BOX, 1, 8, 256, 256)
    photons, passes, seed = 200000, 32, 0x7A75
    r1 = 0.01 * scene.diagonal()
    ctx = P.Context(0)
    ctx.upload(scene)
    ctx.build_bvh()
    ctx.set_camera(scene.camera)
    ctx.begin(photons, passes, r1, seed=seed)
    t0 = time.perf_counter()
    for _ in range(passes):
        ctx.render_pass()
    gpu_s = time.perf_counter() - t0
    g = ctx.image()
    ctx.close()
    os_ = O.OracleScene(scene)
    t0 = time.perf_counter()
    o, _ = os_.render(scene.camera, photons, passes, r1, ray_eps=1e-4, seed=seed)
    cpu_s = time.perf_counter() - t0
    return {
        "workload": "closed box, two f=8 spheres, 256x256, 200k photons/pass, 32 passes",
        "photons": photons * passes,
        "gpu_s": gpu_s,
        "gpu_photons_per_s": photons * passes / gpu_s,
        "cpu_s": cpu_s,
        "cpu_photons_per_s": photons * passes / cpu_s,
        "cpu_threads": int(O.orc().orc_num_threads()),
        "rel_rmse_gpu_vs_fp64_oracle": float(np.sqrt(np.mean((g - o) ** 2)) / o.mean()),
        "note": "single passes (pppm_render_pass) on the GPU; the oracle is FP64 OpenMP",
    }


def oracle_error_study(P, np, lam, probe_every=8):
    """Time to error against the FP64 oracle image of the occluded-light scene
    at 64 x 64 (tests/golden/occluded_reference_64.npz: the oracle's P-KZ, four
    seeds x 256 passes x 1 M photons, written by make_occluded_reference.py),
    plain against learned emission at the same configuration (1 M photons per
    pass, up to 256 passes, GPU seed 0x7A75).

    Error = sqrt(mean((GPU - oracle)^2) - mean(oracle stderr^2)) / mean(oracle):
    the RMSE on linear RGB, the oracle's own noise removed, relative to its
    mean pixel value.  Reports each method's time to relative RMSE 1e-2 (null
    when not reached in 256 passes), and the learned run's time to plain's
    error after 256 passes (speedup = plain time / learned time)."""
    path = os.path.join(ROOT, "tests", "golden", "occluded_reference_64.npz")
    if not os.path.exists(path):
        return {"skipped": f"{path} missing (tests/golden/make_occluded_reference.py writes it)"}
    ref = np.load(path, allow_pickle=False)
    oracle, stderr, cfg = ref["image"], ref["stderr"], ref["config"]
    res, photons, passes = int(cfg[0]), int(cfg[1]), int(cfg[2])
    scene = P.scenegen(P.SCENE_TWO_ROOMS, 1, 100, res, res)
    r1 = 0.01 * scene.diagonal()
    ctx = P.Context(0)
    ctx.upload(scene)
    ctx.build_bvh()
    ctx.set_camera(scene.camera)
    mean_o = float(oracle.mean())
    floor = float(np.mean(stderr ** 2))

    def rel(img):
        return float(np.sqrt(max(float(np.mean((img - oracle) ** 2)) - floor, 0.0))) / mean_o

    def run(learned):
        ctx.begin(photons, passes, r1, seed=0x7A75)
        if learned:
            ctx.proposal_enable(lam=lam, anchors=8, train_interval=8, train_steps=TRAIN[0], batch=TRAIN[1],
                                ring_capacity=16 << 20, learning_rate=2e-3, seed=0x5EED)
        t, curve = 0.0, []
        for i in range(1, passes + 1):
            t0 = time.perf_counter()
            ctx.render_pass()
            t += time.perf_counter() - t0
            if i % probe_every == 0 or i == 1:
                curve.append((i, t, rel(ctx.image())))
        return t, curve

    def time_to(curve, target):
        prev = None
        for i, t, e in curve:
            if e <= target:
                if prev is None:
                    return t
                t0, e0 = prev
                return t0 + (t - t0) * (e0 - target) / max(e0 - e, 1e-30)
            prev = (t, e)
        return None

    t_plain, c_plain = run(False)
    t_learned, c_learned = run(True)
    bound = guiding_bound(ctx, np, photons, r1, 64)
    ctx.close()
    target = c_plain[-1][2]
    reach = time_to(c_learned, target)
    return {
        "scene": "occluded light (two 6^3 rooms, 0.2 x 2 slit), 200k-triangle displaced sphere in room B",
        "resolution": [res, res],
        "photons_per_pass": photons,
        "reference": "FP64 oracle P-KZ, 4 seeds x 256 passes x 1 M photons (tests/golden/occluded_reference_64.npz)",
        "error": "sqrt(MSE - mean oracle stderr^2) / mean(oracle), linear RGB",
        "oracle_rel_noise": float(np.sqrt(floor)) / mean_o,
        "plain_time_to_1e-2_s": time_to(c_plain, 1e-2),
        "learned_time_to_1e-2_s": time_to(c_learned, 1e-2),
        "plain_rel_at_256": target,
        "learned_rel_at_256": c_learned[-1][2],
        "plain_s": t_plain,
        "learned_s_to_plain_error": reach,
        "speedup": (t_plain / reach) if reach else None,
        "learned_total_s": t_learned,
        "lambda": lam,
        "curve_plain": [(i, round(t, 4), round(e, 5)) for i, t, e in c_plain],
        "curve_learned": [(i, round(t, 4), round(e, 5)) for i, t, e in c_learned],
        "emission_guiding_bound": bound,
    }


def guiding_bound(ctx, np, photons, r1, passes):
    """How much an emission-only proposal could lower the variance of the
    image sum on this scene, from the contribution records of plain passes.

    A photon emitted into direction cell k (cosine emission: probability p_k)
    contributes X; E[X^2 | k] = m2_k is estimated from the records (a photon
    with no accepted deposit contributes 0 and leaves no record).  Drawing the
    cell from q and weighting by p_k / q_k gives E[X^2] = sum p_k^2 m2_k / q_k,
    least at q_k ~ p_k sqrt(m2_k) (the zero-variance choice for the first
    segment), so the variance can fall at most by
    (sum p_k m2_k - I^2) / ((sum p_k sqrt(m2_k))^2 - I^2) ("bound", with the
    records' own m2_k, which sparse bins bias upward).  "Cross-validated": q
    built from the even records (mixed with cosine at 0.1) and its variance
    estimated on the odd ones, E[X^2 p / q] over plain samples, over the 32 x
    32 cells and over the cells conditioned on the light's 8 x 8 anchors (the
    learned proposal's table).  Also the records' own target, q ~ p_k E[X | k]
    mixed with cosine at lambda 0.5, which is what the learned proposal fits."""
    ctx.begin(photons, passes, r1, seed=0x7A75 + 17)
    # records only: the retraining interval lies past the run, so emission stays cosine
    ctx.proposal_enable(lam=0.5, anchors=8, train_interval=1 << 30, train_steps=1, batch=256,
                        ring_capacity=1 << 26, learning_rate=0.0, seed=1)
    for _ in range(passes):
        ctx.render_pass()
    rec = ctx.proposal_records()
    n = float(photons) * passes
    cell = rec[:, 2].astype(np.int64)
    c = rec[:, 3].astype(np.float64)
    side = 32
    cos_hi = 1.0 - np.arange(side) / side
    cos_lo = 1.0 - (np.arange(side) + 1) / side
    p = np.repeat((cos_hi ** 2 - cos_lo ** 2) / side, side)  # sin^2(theta) uniform, phi uniform
    u = np.minimum((rec[:, 0] * 8).astype(np.int64), 7)
    v = np.minimum((rec[:, 1] * 8).astype(np.int64), 7)
    anchor_bins, anchor_p = (u * 8 + v) * 1024 + cell, np.tile(p / 64.0, 64)

    s2 = np.bincount(cell, weights=c * c, minlength=1024)
    e2 = s2.sum() / n
    i2 = (c.sum() / n) ** 2
    plug_in = (e2 - i2) / (np.sum(np.sqrt(p * s2 / n)) ** 2 - i2)

    def cross_validated(bins, pb, eps=0.1):
        even = np.arange(len(c)) % 2 == 0
        qa = np.sqrt(pb * np.bincount(bins[even], weights=c[even] ** 2, minlength=len(pb)))
        q = (1.0 - eps) * qa / qa.sum() + eps * pb
        odd = ~even
        half = n / 2.0
        e2b = np.sum(c[odd] ** 2) / half
        e2q = np.sum(c[odd] ** 2 * pb[bins[odd]] / q[bins[odd]]) / half
        i2b = (np.sum(c[odd]) / half) ** 2
        return (e2b - i2b) / (e2q - i2b)

    s1 = np.bincount(cell, weights=c, minlength=1024)
    q = 0.5 * p + 0.5 * s1 / s1.sum()
    v_mle = np.sum(p * p * (s2 / (n * p)) / q)
    return {
        "passes": passes,
        "photons_per_pass": photons,
        "contributing_fraction": len(rec) / n,
        "variance_reduction_bound": plug_in,
        "variance_reduction_cross_validated": cross_validated(cell, p),
        "variance_reduction_cross_validated_per_anchor": cross_validated(anchor_bins, anchor_p),
        "mle_target_lambda_0.5_reduction": (e2 - i2) / (v_mle - i2),
        "note": "first-segment variance of the image sum under emission-only proposals, from plain-emission "
                "contribution records: no proposal over the 32 x 32 cells lowers it by more than the bound",
    }


# Dipole sum issue model (gfx950, wave64): the inner loop of k_dipole issues
# DIPOLE_VALU full-rate and DIPOLE_TRANS transcendental (v_rsq_f32, v_exp_f32)
# VALU instructions per pair (counted in its AMD assembly); their cycle costs
# come from pppm_issue_probe, measured in the same run.
DIPOLE_VALU = 65
DIPOLE_TRANS = 12


def dipole_peaks(cus, probe):
    lanes_per_s = cus * 4 * probe["clock_hz"] * 64
    issue = lanes_per_s / (probe["valu_cycles"] * DIPOLE_VALU + probe["trans_cycles"] * DIPOLE_TRANS)
    trans = lanes_per_s / (probe["trans_cycles"] * DIPOLE_TRANS)
    return issue, trans


def subsurface_study(P, np, width, photons, passes, points, frequency):
    """The translucent two-room scene: displaced icosphere (frequency f gives
    20 f^2 triangles) with a dipole BSSRDF, `points` surface points, the
    learned proposal on.  Reports the per-pass stages and the dipole sum's
    pair evaluations per second against the issue model."""
    scene = P.scenegen(P.SCENE_TWO_ROOMS_SSS, 1, frequency, width, width)
    r1 = 0.01 * scene.diagonal()
    ctx = P.Context(0)
    cus = ctx.device_info()["cu"]
    probe = ctx.issue_probe()
    ctx.upload(scene)
    bvh = ctx.build_bvh()
    ctx.set_camera(scene.camera)
    ctx.subsurface_configure(points)

    def run(max_error):
        ctx.subsurface_cut(max_error)
        t0 = time.perf_counter()
        ctx.begin(photons, passes + 1, r1, seed=0x7A75)
        setup = time.perf_counter() - t0
        ctx.proposal_enable(lam=0.5, anchors=8, train_interval=8, train_steps=TRAIN[0], batch=TRAIN[1],
                            ring_capacity=16 << 20, learning_rate=2e-3, seed=0x5EED)
        ctx.render_pass()  # warm-up
        out = []
        t0 = time.perf_counter()
        for _ in range(passes):
            out.append(ctx.render_pass())
        return out, time.perf_counter() - t0, setup, ctx.image()

    sts, wall, setup_s, img = run(0.0)
    # the reference's octree cut (max_error 0.05): same photons and eye paths
    sts_c, wall_c, setup_c, img_c = run(0.05)
    ctx.close()
    assert np.all(np.isfinite(img)) and np.all(np.isfinite(img_c))
    pairs = sum(s["sss_hits"] * s["sss_points"] for s in sts)
    dip_s = sum(s["dipole_ms"] for s in sts) * 1e-3
    issue, trans = dipole_peaks(cus, probe)
    achieved = pairs / dip_s if dip_s > 0 else 0.0
    return {
        "workload": "two rooms, translucent displaced icosphere (dipole BSSRDF), learned proposal on",
        "triangles": int(scene.n_tris),
        "bvh_build_ms": bvh["build_ms"],
        "resolution": [width, width],
        "photons_per_pass": photons,
        "passes": passes,
        "surface_points": int(sts[0]["sss_points"]),
        "points_setup_s": setup_s,
        "ms_per_pass": 1e3 * wall / passes,
        "photons_per_s": photons * passes / wall,
        "stage_ms": {k: sum(s[k] for s in sts) / passes
                     for k in ("photon_ms", "irradiance_ms", "order_ms", "grid_ms", "eye_ms", "gather_ms",
                               "proposal_ms", "dipole_ms", "total_ms")},
        "irradiance_photons_per_pass": sum(s["irradiance_photons"] for s in sts) / passes,
        "sss_hits_per_pass": sum(s["sss_hits"] for s in sts) / passes,
        "octree_cut": {
            "max_error": 0.05,
            "ms_per_pass": 1e3 * wall_c / passes,
            "dipole_ms": sum(s["dipole_ms"] for s in sts_c) / passes,
            "setup_s": setup_c,
            "image_rel_l1_vs_exact": float(np.abs(img_c - img).sum() / img.sum()),
            "note": "k_dipole_cut plus the per-level node updates; the exact all-pairs sum stays the parity path",
        },
        "dipole_roofline": {
            "bound": "valu",
            "kernel": "k_dipole",
            "achieved": achieved,
            "peak": issue,
            "unit": "pair evaluations/s",
            "frac": achieved / issue,
            "transcendental_peak": trans,
            "frac_of_transcendental_peak": achieved / trans,
            "model": (f"{DIPOLE_VALU} full-rate VALU x {probe['valu_cycles']:.2f} cycles + {DIPOLE_TRANS} "
                      f"transcendental x {probe['trans_cycles']:.2f} cycles per pair and wave64 (measured), "
                      f"{cus} CUs x 4 SIMDs at {probe['clock_hz'] / 1e9:.2f} GHz"),
        },
    }


def batch_study(P, np, n_scenes, passes, seed, branches=4):
    """Batched plain PPM (config C3): n_scenes procedural scenes, triangle
    counts log-uniform in [1k, 500k], resolutions from {128, 256, 512, 1024}^2,
    8 photons per pixel, in one hipGraph-captured pass loop after the
    autotuner; against the same scenes rendered one after another with
    pppm_render_pass."""
    specs = P.batch_specs(n_scenes, seed)
    t0 = time.perf_counter()
    b = P.Batch(branches=branches)
    scenes = []
    for k, (kind, sd, f, res) in enumerate(specs):
        s = P.scenegen(kind, sd, f, res, res)
        n = 8 * res * res
        r1 = 0.01 * s.diagonal()
        b.add(s, n, passes + 1, r1, hash_size=n, seed=0x900 + k, capacity_factor=4.0)
        scenes.append((s, n, r1, 0x900 + k))
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.tune()
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.run(1)  # captures the graph and warms up
    capture_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    gpu_ms = b.run(passes)
    wall = time.perf_counter() - t0
    infos = [b.info(k) for k in range(n_scenes)]
    for k in (0, n_scenes - 1):
        assert np.all(np.isfinite(b.image(k)))
    b.close()
    photons = sum(n for _, n, _, _ in scenes)
    # one scene at a time through pppm_render_pass, one warm-up and one timed pass
    ctx = P.Context(0)
    solo_ms = 0.0
    for k, (s, n, r1, sd) in enumerate(scenes):
        ctx.upload(s)
        ctx.build_bvh()
        ctx.set_camera(s.camera)
        ctx.begin(n, 2, r1, seed=sd, hash_size=n, gather_mode=infos[k]["gather_mode"], capacity_factor=4.0)
        ctx.render_pass()
        t1 = time.perf_counter()
        ctx.render_pass()
        solo_ms += 1e3 * (time.perf_counter() - t1)
    cu = max(1, ctx.device_info()["cu"])
    ctx.close()
    modes = {P.GATHER_LANES: "lanes", P.GATHER_TILED: "tiled", P.GATHER_WAVE: "wave", P.GATHER_BLOCKS: "blocks"}
    return {
        "workload": (f"{n_scenes} scenes, triangles log-uniform in [1k, 500k], resolutions from "
                     "{128,256,512,1024}^2, 8 photons per pixel"),
        "branches": branches,
        "passes_timed": passes,
        "photons_per_batch_pass": photons,
        "triangles_total": int(sum(s.n_tris for s, _, _, _ in scenes)),
        "pixels_total": int(sum(s.camera.width * s.camera.height for s, _, _, _ in scenes)),
        "ms_per_batch_pass": 1e3 * wall / passes,
        "gpu_ms_per_batch_pass": gpu_ms / passes,
        "photons_per_s": photons * passes / wall,
        "solo_ms_per_batch_pass": solo_ms,
        "speedup_vs_solo": solo_ms / (1e3 * wall / passes),
        "setup_s": setup_s,
        "tune_s": tune_s,
        "capture_and_first_pass_s": capture_s,
        "tuned_gather": {v: sum(1 for i in infos if i["gather_mode"] == k) for k, v in modes.items()},
        "tuned_photon_blocks_per_cu": {str(g): sum(1 for i in infos if i["photon_grid"] // cu == g)
                                       for g in sorted({i["photon_grid"] // cu for i in infos})},
        "tuned_trace_keep": {str(k): sum(1 for i in infos if i["trace_keep"] == k)
                             for k in sorted({i["trace_keep"] for i in infos})},
    }


def whole_job(ctx, photons, passes, r1, ray_eps, seed, mode):
    """The whole progressive job in one pipelined call, from pass 1: whole-job
    photons/s (pipeline fill and drain included) and the gather's useful-byte
    rate over ranges of passes. The radius shrinks as r_i^2 = r_1^2 prod_k
    (k + alpha)/(k + 1), so the in-radius photons (the useful bytes) fall with
    it while the per-hit-point work stays: the gather's HBM fraction falls
    over the job."""
    t0 = time.perf_counter()
    ctx.begin(photons, passes, r1, ray_eps=ray_eps, seed=seed, gather_mode=mode)
    begin_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    stats = ctx.render_passes(passes)
    elapsed = time.perf_counter() - t0
    edges = [b for b in (1, 3, 11, 33, 65, 129) if b <= passes] + [passes + 1]
    ranges = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        part = stats[lo - 1:hi - 1]
        g_ms = sum(s["gather_ms"] for s in part) / len(part)
        useful = sum(useful_bytes(s) for s in part) / len(part)
        ranges.append({"passes": [lo, hi - 1], "gather_avg_launch_ms": g_ms,
                       "useful_bytes_per_launch": useful,
                       "accepted_per_hitpoint": sum(s["gather_inradius"] for s in part)
                       / max(1, sum(s["hitpoints"] for s in part)),
                       "frac": useful / (g_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS})
    g_ms = sum(s["gather_ms"] for s in stats)
    useful = sum(useful_bytes(s) for s in stats)
    return {"passes": passes, "photons_per_pass": photons, "begin_s": begin_s, "wall_s": elapsed,
            "photons_per_s": photons * passes / elapsed, "ms_per_pass": 1e3 * elapsed / passes,
            "gather_frac": useful / (g_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
            "gather_by_passes": ranges,
            "note": "one render_passes call over the whole job from pass 1 (fill and drain included); "
                    "gather fraction = useful bytes / pipelined launch span / HBM peak"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--job-passes", type=int, default=256,
                    help="passes of the whole-job leg (0 skips it): one pipelined run from pass 1")
    ap.add_argument("--photons", type=int, default=4_000_000)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--gather-mode", choices=["auto", "lanes", "tiled", "wave", "blocks"], default="auto")
    ap.add_argument("--bvh", choices=["default", "lbvh", "ploc"], default="default",
                    help="binary tree the 4-wide BVH of the headline scene is collapsed from")
    ap.add_argument("--no-rmse", action="store_true", help="skip the plain-vs-learned time-to-error leg")
    ap.add_argument("--rmse-passes", type=int, default=256)
    ap.add_argument("--rmse-ref-passes", type=int, default=1024)
    ap.add_argument("--rmse-lambda", type=float, default=0.5)
    ap.add_argument("--train-steps", type=int, default=8, help="Adam steps per retrain of the proposal MLP")
    ap.add_argument("--train-batch", type=int, default=65536, help="samples per step (a multiple of 256)")
    ap.add_argument("--no-sss", action="store_true", help="skip the subsurface (dipole BSSRDF) leg")
    ap.add_argument("--sss-passes", type=int, default=2)
    ap.add_argument("--sss-photons", type=int, default=8_000_000)
    ap.add_argument("--sss-points", type=int, default=1 << 20)
    ap.add_argument("--sss-frequency", type=int, default=316)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-passes", type=int, default=64)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="after the passes, copy BYTES device-to-device once (a plain copy kernel "
                         "of known traffic that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    TRAIN[0], TRAIN[1] = args.train_steps, args.train_batch
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
    bvh = ctx.build_bvh({"default": None, "lbvh": P.BVH_LBVH, "ploc": P.BVH_PLOC}[args.bvh])
    ctx.set_camera(scene.camera)
    total = args.warmup + args.steps
    mode = {"auto": P.GATHER_AUTO, "lanes": P.GATHER_LANES, "tiled": P.GATHER_TILED,
            "wave": P.GATHER_WAVE, "blocks": P.GATHER_BLOCKS}[args.gather_mode]
    ctx.begin(args.photons, total, r1, ray_eps=ray_eps, seed=seed, gather_mode=mode)
    for _ in range(args.warmup):
        ctx.render_pass()
    t0 = time.perf_counter()
    # pipelined: the gather of pass i overlaps the photons of pass i + 1;
    # returns after the last pass completes
    stats = ctx.render_passes(args.steps)
    elapsed = time.perf_counter() - t0
    img = ctx.image()
    assert np.all(np.isfinite(img))
    job = whole_job(ctx, args.photons, args.job_passes, r1, ray_eps, seed + 2, mode) if args.job_passes > 0 else None
    # the gather alone (single passes: nothing runs beside it), untimed, over
    # the timed pass range (same radii); these passes also count BVH nodes and
    # triangles per ray (counting kernels, same images), which the gather does
    # not depend on
    ctx.begin(args.photons, total, r1, ray_eps=ray_eps, seed=seed + 1, gather_mode=mode)
    for _ in range(args.warmup):
        ctx.render_pass()
    ctx.count_traversal(True)
    iso = [ctx.render_pass() for _ in range(args.steps)]
    ctx.count_traversal(False)
    probe = ctx.issue_probe()
    cus = max(1, ctx.device_info()["cu"])
    iso_ms = sum(s["gather_ms"] for s in iso) / len(iso)
    iso_photon_ms = sum(s["photon_ms"] for s in iso) / len(iso)
    iso_bytes = sum(gather_bytes(s) for s in iso) / len(iso)
    iso_useful = sum(useful_bytes(s) for s in iso) / len(iso)
    rays_c = sum(s["photon_segments"] for s in iso)
    nodes_per_ray = sum(s["photon_nodes"] for s in iso) / rays_c
    tris_per_ray = sum(s["photon_tris"] for s in iso) / rays_c
    eye_rays_c = sum(s["eye_segments"] for s in iso)
    eye_nodes_per_ray = sum(s["eye_nodes"] for s in iso) / eye_rays_c
    eye_tris_per_ray = sum(s["eye_tris"] for s in iso) / eye_rays_c

    photons_per_s = args.photons * args.steps / elapsed
    segs = sum(s["photon_segments"] + s["eye_segments"] for s in stats)
    gather_ms = sum(s["gather_ms"] for s in stats) / args.steps
    gbytes = sum(gather_bytes(s) for s in stats) / args.steps
    ref_cell_rate = gbytes / (gather_ms * 1e-3) / 1e9
    photon_kernel_ms = sum(s["photon_ms"] for s in stats) / args.steps
    photon_rays = sum(s["photon_segments"] for s in stats) / args.steps
    trav_bytes = traversal_bytes(photon_rays, nodes_per_ray, tris_per_ray)
    trav_gbps = trav_bytes / (photon_kernel_ms * 1e-3) / 1e9
    useful = sum(useful_bytes(s) for s in stats) / args.steps
    achieved = useful / (gather_ms * 1e-3) / 1e9
    tested = sum(s["gather_tested"] for s in stats) / args.steps
    hits = sum(s["hitpoints"] for s in stats) / args.steps
    read_bytes = hits * (HIT_RECORD_BYTES + ACCUM_BYTES) + PHOTON_BYTES * tested
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
            "bvh": {**{k: bvh[k] for k in ("nodes", "leaves", "depth", "build_ms")},
                    "builder": {P.BVH_LBVH: "lbvh", P.BVH_PLOC: "ploc"}.get(bvh["builder"], "?")},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_gather_lanes",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": read_traffic(),
            "algorithmic_bytes_per_launch": useful,
            "bytes_formula": "H (40 + 48) + 24 per in-radius photon: hit point record, Kahan pair read and "
                             "written, position and flux of every photon the estimate sums",
            "avg_launch_ms": gather_ms,
            "note": "timed region runs passes pipelined: the gather shares the GPU with the next pass's photons",
            "isolated": {"passes": [args.warmup + 1, total],
                         "achieved": iso_useful / (iso_ms * 1e-3) / 1e9, "avg_launch_ms": iso_ms,
                         "frac": iso_useful / (iso_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS},
            "records_read_rate": {"bytes_per_launch": read_bytes,
                                  "achieved": read_bytes / (gather_ms * 1e-3) / 1e9,
                                  "note": "H (40 + 48) + 24 per photon record the kernel distance-tests"},
            "reference_cell_equivalent": {
                "bytes_per_launch": gbytes, "rate": ref_cell_rate,
                "isolated_rate": iso_bytes / (iso_ms * 1e-3) / 1e9,
                "note": "H (40 + 48) + 8 per bucket + 24 per record of the buckets of the reference's 2r "
                        "cells, whether read or not; a work-equivalent rate, not bandwidth"},
            "issue": issue_bound("k_gather_lanes", probe, cus, iso_ms, gather_ms),
            "records_tested_per_launch": tested,
            "records_in_buckets_per_launch": sum(s["gather_candidates"] for s in stats) / args.steps,
            "records_accepted_per_launch": sum(s["gather_inradius"] for s in stats) / args.steps,
        },
        "traversal": {
            "kernel": "k_photon",
            "rays_per_launch": photon_rays,
            "nodes_per_ray": nodes_per_ray,
            "tris_per_ray": tris_per_ray,
            "eye_nodes_per_ray": eye_nodes_per_ray,
            "eye_tris_per_ray": eye_tris_per_ray,
            "algorithmic_bytes_per_launch": trav_bytes,
            "traffic": read_traffic("photon_pmc_caustic_box_1024.txt"),
            "avg_launch_ms": photon_kernel_ms,
            "achieved": trav_gbps,
            "unit": "GB/s",
            "frac": trav_gbps / HBM_PEAK_GBPS,
            "rays_per_s": photon_rays / (photon_kernel_ms * 1e-3),
            "issue": issue_bound("k_photon<true, false>", probe, cus, iso_photon_ms, None),
            "note": "bytes = rays (32 + 16) + 128 per node + 36 per triangle; node and triangle counts from "
                    "the counting kernels on untimed passes; the BVH (about 50 MB) lives in the 256 MB "
                    "last-level cache",
        },
        "rays_per_s": segs / elapsed,
        "photon_kernel_photons_per_s": args.photons / (photon_kernel_ms * 1e-3),
        "gather_gbps": achieved,
        "gather_reference_cell_equivalent_gbps": ref_cell_rate,
        "stage_span_ms": {k: sum(s[k] for s in stats) / args.steps
                          for k in ("photon_ms", "order_ms", "grid_ms", "eye_ms", "hitsort_ms", "gather_ms",
                                    "total_ms")},
        "stage_span_note": ("HIP-event spans on four concurrent streams, averaged over the timed passes: each "
                            "includes the time its kernel shares the chip with the others, so they overlap and "
                            "sum past ms_per_step; kernel durations are in profiles/kernels_caustic_box_1024.txt"),
        "gather_mode": args.gather_mode,
        "photons_stored_per_pass": sum(s["photons_stored"] for s in stats) / args.steps,
        "hitpoints_per_pass": sum(s["hitpoints"] for s in stats) / args.steps,
        "whole_job": job,
        "time_to_rmse": None,
        "time_to_rmse_vs_fp64_oracle": None,
    }
    if args.copy_calibration > 0:
        copy_ms = ctx.stream_copy(args.copy_calibration)
        out["copy_calibration"] = {"bytes": args.copy_calibration, "ms": copy_ms,
                                   "GBps": 2 * args.copy_calibration / (copy_ms * 1e-3) / 1e9}
    ctx.close()
    if not args.no_rmse:
        out["time_to_rmse"] = error_study(P, np, args.photons, args.width, args.rmse_passes,
                                          args.rmse_ref_passes, args.rmse_lambda)
        out["time_to_rmse_vs_fp64_oracle"] = oracle_error_study(P, np, args.rmse_lambda)
    if not args.no_sss:
        out["subsurface"] = subsurface_study(P, np, args.width, args.sss_photons, args.sss_passes,
                                             args.sss_points, args.sss_frequency)
    else:
        out["subsurface"] = None
    if not args.no_batch:
        out["batch"] = batch_study(P, np, args.batch_scenes, args.batch_passes, 7)
    else:
        out["batch"] = None
    if not args.no_cpu:
        cb, _ = cpu_baseline(scene, args.photons, r1, ray_eps, seed)
        cb["c1_full_run"] = c1_parity_run(P, np)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
