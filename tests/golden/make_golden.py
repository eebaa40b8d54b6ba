"""Writes tests/golden/reference_vectors.npz from the compiled tatsy-pppm
reference (oracle/_ref/libref.so, built by `make -C oracle ref` from the
reference's own sources).

Every array in the file is data: seeded inputs and the reference's outputs for
them.  Functions covered: HashGrid<int> add/operator[] (hash_grid_detail.h),
QBVHAccel construct/intersect (qbvh_accel.cc), Triangle::intersect
(triangle.cc), sampler::onHemisphere (sampler.cc), checkTotalReflection
(reflectance.h), BSDF::sample of the Lambertian / specular / refraction BSDFs
(brdf.cc), PerspectiveCamera::getRay (perspective_camera.cc) and
DipoleBSSRDF::operator() / Fdr (bssrdf.cc) and SubsurfaceIntegrator's octree
cut of the dipole sum (subsurface_integrator.cc:69-167, 234-238).

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

P = oracle._ptr
OUT = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")


def hash_cases(rng):
    """(name, positions, hash_size, scale, bbox_min) point-insertion cases."""
    cases = []
    pts = rng.uniform(0.0, 5.0, (3000, 3)).astype(np.float32)
    cases.append(("uniform", pts, 2999, 1.0 / (2 * 0.05), np.zeros(3)))
    # points exactly on cell boundaries (cell edge 0.25, coordinates multiples of it)
    grid = (rng.integers(0, 40, (2000, 3)) * 0.25).astype(np.float32)
    cases.append(("boundaries", grid, 1021, 4.0, np.zeros(3)))
    # coordinates far from the origin: ix * 73856093 wraps in int32
    far = rng.uniform(0.0, 3.0, (2000, 3)).astype(np.float32)
    cases.append(("wrap", far, 4096 + 7, 1.0 / (2 * 0.001), np.zeros(3)))
    # negative coordinates relative to bbox_min: abs() folding
    neg = rng.uniform(-2.0, 2.0, (2000, 3)).astype(np.float32)
    cases.append(("negative", neg, 997, 3.0, np.zeros(3)))
    # duplicates and a tiny table (forced collisions)
    dup = np.repeat(rng.uniform(0.0, 1.0, (100, 3)).astype(np.float32), 5, axis=0)
    rng.shuffle(dup)
    cases.append(("duplicates", dup, 7, 10.0, np.zeros(3)))
    cases.append(("single", np.array([[0.3, 0.2, 0.1]], np.float32), 13, 2.0, np.zeros(3)))
    return cases


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libref.so is missing: run `make -C oracle ref` where /root/reference exists")
    R = oracle.ref()
    rng = np.random.default_rng(20261202)
    out = {}

    # --- HashGrid<int>, point insertion: bucket list of every point -------
    for name, pos, S, scale, bmin in hash_cases(rng):
        n = len(pos)
        cap = n * n + 16
        lists = np.zeros(cap, np.int32)
        off = np.zeros(n + 1, np.int64)
        bmin = np.asarray(bmin, np.float64)
        bmax = bmin + 1.0
        k = R.ref_hashgrid_point_buckets(P(pos), n, S, scale, P(bmin), P(bmax), P(lists), cap, P(off))
        assert k >= 0
        out[f"hg_{name}_pos"] = pos
        out[f"hg_{name}_params"] = np.array([S, scale, *bmin], np.float64)
        out[f"hg_{name}_lists"] = lists[:k].copy()
        out[f"hg_{name}_off"] = off

    # --- HashGrid<int>, box insertion (Hachisuka-style render points) ------
    nb = 400
    centers = rng.uniform(0.2, 4.8, (nb, 3))
    rad = 0.11
    bmin_i = centers - rad
    bmax_i = centers + rad
    queries = np.concatenate([centers, rng.uniform(0.0, 5.0, (300, 3))])
    S, scale, gmin = 797, 1.0 / (2 * rad), np.zeros(3)
    cap = 400 * 400 * 8
    lists = np.zeros(cap, np.int32)
    off = np.zeros(len(queries) + 1, np.int64)
    k = R.ref_hashgrid_box_buckets(P(bmin_i), P(bmax_i), nb, S, scale, P(gmin), P(gmin + 1.0), P(queries),
                                   len(queries), P(lists), cap, P(off))
    assert k >= 0
    out["hgbox_bmin"] = bmin_i
    out["hgbox_bmax"] = bmax_i
    out["hgbox_queries"] = queries
    out["hgbox_params"] = np.array([S, scale, *gmin], np.float64)
    out["hgbox_lists"] = lists[:k].copy()
    out["hgbox_off"] = off

    # --- QBVHAccel on a random triangle soup --------------------------------
    nt = 500
    c = rng.uniform(0.0, 4.0, (nt, 1, 3))
    tri = (c + rng.normal(scale=0.3, size=(nt, 3, 3))).reshape(nt, 9).astype(np.float32)
    nr = 3000
    rays = np.zeros((nr, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-0.5, 4.5, (nr, 3))
    d = rng.normal(size=(nr, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    rays = np.ascontiguousarray(rays)
    # the Ray constructor asserts a unit direction in FP64: renormalise the FP32 copy
    dd = rays[:, 4:7].astype(np.float64)
    assert np.all(np.abs(1.0 - np.linalg.norm(dd, axis=1)) < 1e-3)
    t = np.zeros(nr, np.float64)
    ids = np.zeros(nr, np.int32)
    R.ref_qbvh_intersect(P(tri), nt, P(rays), nr, P(t), P(ids))
    out["qbvh_tri"] = tri
    out["qbvh_rays"] = rays
    out["qbvh_t"] = t
    out["qbvh_id"] = ids

    # --- Triangle::intersect -----------------------------------------------
    m = 2000
    tt = rng.normal(size=(m, 9))
    oo = rng.normal(size=(m, 3)) * 2.0
    tgt = tt[:, 0:3] * 0.3 + tt[:, 3:6] * 0.3 + tt[:, 6:9] * 0.4 + rng.normal(scale=0.3, size=(m, 3))
    dd = tgt - oo
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    tri_hit = np.zeros(m, np.int32)
    tri_t = np.zeros(m)
    tri_n = np.zeros((m, 3))
    for i in range(m):
        tv = np.ascontiguousarray(tt[i]); ov = np.ascontiguousarray(oo[i]); dv = np.ascontiguousarray(dd[i])
        t1 = C.c_double()
        nv = np.zeros(3)
        tri_hit[i] = R.ref_triangle_intersect(P(tv), P(ov), P(dv), C.byref(t1), P(nv))
        tri_t[i] = t1.value
        tri_n[i] = nv
    out["tri_tri"], out["tri_o"], out["tri_d"] = tt, oo, dd
    out["tri_hit"], out["tri_t"], out["tri_n"] = tri_hit, tri_t, tri_n

    # --- sampler::onHemisphere ---------------------------------------------
    m = 2000
    nrm = rng.normal(size=(m, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[:50] = [0.0, -1.0, 0.0]  # |w.x| <= EPS branch
    nrm[50:100] = [0.0, 0.0, 1.0]
    r12 = rng.uniform(0.0, 1.0, (m, 2))
    hemi = np.zeros((m, 3))
    for i in range(m):
        o3 = np.zeros(3)
        R.ref_on_hemisphere(P(np.ascontiguousarray(nrm[i])), r12[i, 0], r12[i, 1], P(o3))
        hemi[i] = o3
    out["hemi_n"], out["hemi_r"], out["hemi_out"] = nrm, r12, hemi

    # --- checkTotalReflection and BSDF::sample -----------------------------
    m = 1500
    din = rng.normal(size=(m, 3))
    din /= np.linalg.norm(din, axis=1, keepdims=True)
    nn = rng.normal(size=(m, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    into = (np.einsum("ij,ij->i", din, nn) < 0).astype(np.int32)
    on = np.where(into[:, None] == 1, nn, -nn)
    tr_flag = np.zeros(m, np.int32)
    tr_vals = np.zeros((m, 8))
    for i in range(m):
        rf, rr = np.zeros(3), np.zeros(3)
        fre, ftr = C.c_double(), C.c_double()
        tr_flag[i] = R.ref_total_reflection(int(into[i]), P(np.ascontiguousarray(din[i])),
                                            P(np.ascontiguousarray(nn[i])), P(np.ascontiguousarray(on[i])),
                                            P(rf), P(rr), C.byref(fre), C.byref(ftr))
        tr_vals[i] = [*rf, *rr, fre.value, ftr.value]
    out["fr_in"], out["fr_n"], out["fr_into"], out["fr_on"] = din, nn, into, on
    out["fr_flag"], out["fr_vals"] = tr_flag, tr_vals

    types = np.array([0x01, 0x02, 0x08], np.int32)
    rho = np.array([0.7, 0.5, 0.3])
    bs_out = np.zeros((3, m, 4))
    r12 = rng.uniform(0.0, 1.0, (m, 2))
    for k, ty in enumerate(types):
        for i in range(m):
            o3 = np.zeros(3)
            pdf = C.c_double(1.0)
            R.ref_bsdf_sample(int(ty), P(rho), P(np.ascontiguousarray(din[i])), P(np.ascontiguousarray(nn[i])),
                              r12[i, 0], r12[i, 1], P(o3), C.byref(pdf))
            bs_out[k, i] = [*o3, pdf.value]
    out["bs_types"], out["bs_rho"], out["bs_r"], out["bs_out"] = types, rho, r12, bs_out

    # --- PerspectiveCamera::getRay -----------------------------------------
    cams = []
    for center, direc, up, fov, w, h in [((2.5, 2.5, 4.9), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 1.0, 256, 256),
                                          ((0.0, 1.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 0.8, 320, 200)]:
        pxy = np.concatenate([rng.uniform(0, w, (200, 1)), rng.uniform(0, h, (200, 1))], axis=1)
        pxy[0] = [w / 2, h / 2]
        res = np.zeros((200, 6))
        for i in range(200):
            o3, d3 = np.zeros(3), np.zeros(3)
            R.ref_camera_ray(P(np.array(center)), P(np.array(direc)), P(np.array(up)), fov, w, h,
                             pxy[i, 0], pxy[i, 1], P(o3), P(d3))
            res[i] = [*o3, *d3]
        cams.append(np.array([*center, *direc, *up, fov, w, h], np.float64))
        out[f"cam{len(cams) - 1}_params"] = cams[-1]
        out[f"cam{len(cams) - 1}_pxy"] = pxy
        out[f"cam{len(cams) - 1}_rays"] = res

    # --- DipoleBSSRDF -------------------------------------------------------
    dip_params = []
    dip_vals = []
    for _ in range(20):
        sa = rng.uniform(0.001, 0.05, 3)
        ss = rng.uniform(0.5, 2.0, 3)
        eta = 1.3
        d2s = np.concatenate([[0.0], rng.uniform(0.0, 4.0, 15)])
        for d2 in d2s:
            o3 = np.zeros(3)
            fdr = C.c_double()
            R.ref_dipole(P(sa), P(ss), eta, d2, P(o3), C.byref(fdr))
            dip_params.append([*sa, *ss, eta, d2])
            dip_vals.append([*o3, fdr.value])
    out["dip_params"] = np.array(dip_params)
    out["dip_vals"] = np.array(dip_vals)

    # --- SubsurfaceIntegrator's octree cut (its own generator, so the
    # arrays above do not move) ------------------------------------------------
    orng = np.random.default_rng(0x0C7)
    for name, n, eps in (("sphere", 3000, 0.05), ("plane", 2500, 0.05), ("sphere_tight", 3000, 0.01)):
        if name.startswith("sphere"):
            v = orng.normal(size=(n, 3))
            pts = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
            qv = orng.normal(size=(200, 3))
            q = (qv / np.linalg.norm(qv, axis=1, keepdims=True) * orng.uniform(0.97, 1.03, (200, 1)))
        else:
            pts = np.column_stack([orng.uniform(-1.5, 1.5, n), np.zeros(n), orng.uniform(-1.5, 1.5, n)])
            q = np.column_stack([orng.uniform(-2.0, 2.0, 200), orng.uniform(-0.05, 0.05, 200),
                                 orng.uniform(-2.0, 2.0, 200)])
        pts = np.ascontiguousarray(pts, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        E = np.ascontiguousarray(orng.uniform(0.0, 2.0, (n, 3)) * (orng.random((n, 1)) < 0.9), np.float32)
        sa, ss, eta = orng.uniform(0.01, 0.2, 3), orng.uniform(0.5, 3.0, 3), 1.3
        area = 4.0 * np.pi / n if name.startswith("sphere") else 9.0 / n
        res = np.zeros((len(q), 3))
        R.ref_dipole_octree(P(sa), P(ss), eta, P(pts), P(E), n, area, P(q), len(q), eps, P(res))
        out[f"oct_{name}_pts"] = pts
        out[f"oct_{name}_E"] = E
        out[f"oct_{name}_q"] = q
        out[f"oct_{name}_params"] = np.array([*sa, *ss, eta, area, eps])
        out[f"oct_{name}_out"] = res

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
