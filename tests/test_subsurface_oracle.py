"""CPU checks that pin the dipole BSSRDF of the oracle.

The oracle restates DipoleBSSRDF (sources/bssrdf.cc:33-104).  It is pinned by
the compiled reference's own values (golden vectors written with the
reference's Schraudolph exponential), and, for the true exponential the build
uses, by the closed form of the profile's area integral: with
d = sqrt(r^2 + z^2), d/dd(-e^{-s d} / d) = (s d + 1) e^{-s d} / d^2, so each
source term integrates over r in [0, r*] in closed form, r* being where the
clamped profile reaches 0.
"""
import math

import numpy as np
import pytest

import oracle as O

MARBLE = ((0.0021, 0.0041, 0.0071), (2.19, 2.62, 3.00))  # sources/main/appmain.cc:44-45


def test_dipole_matches_reference_golden(golden):
    prm, val = golden["dip_params"], golden["dip_vals"]
    for p, v in zip(prm, val):
        r = O.dipole(p[0:3], p[3:6], p[6], p[7], fast=True)
        np.testing.assert_allclose(r, v[:3], rtol=1e-12, atol=1e-300)
        assert O.orc().orc_dipole_fdr(p[6]) == pytest.approx(v[3], rel=1e-14)


def test_fdr_fit_values():
    assert abs(O.orc().orc_dipole_fdr(1.0)) < 3e-3            # no boundary: ~ no internal reflection
    assert O.orc().orc_dipole_fdr(1.3) == pytest.approx(0.4453776, rel=1e-6)
    assert 0.0 < O.orc().orc_dipole_fdr(1.5) < 1.0


def _closed_integral(sa, ss, eta, c):
    F = O.orc().orc_dipole_fdr(eta)
    A = (1 + F) / (1 - F)
    st = sa[c] + ss[c]
    s = math.sqrt(3 * sa[c] * st)
    ap, zp = ss[c] / st, 1 / st
    zn = zp * (1 + 4 / 3 * A)
    lo, hi = 0.0, 100.0   # r*: bisection on the oracle's own sign change
    for _ in range(200):
        m = 0.5 * (lo + hi)
        if O.dipole(sa, ss, eta, m * m)[c] > 0:
            lo = m
        else:
            hi = m
    dp, dn = math.hypot(lo, zp), math.hypot(lo, zn)
    return ap / (2 * s) * (math.exp(-s * zp) - zp * math.exp(-s * dp) / dp - math.exp(-s * zn) + zn * math.exp(-s * dn) / dn)


@pytest.mark.parametrize("sa,ss", [MARBLE, ((0.01, 0.02, 0.04), (1.0, 1.5, 0.7))])
def test_area_integral_closed_form(sa, ss):
    rs = np.linspace(0.0, 30.0, 300001)
    R = np.array([O.dipole(sa, ss, 1.3, r * r) for r in rs[::10]])
    for c in range(3):
        # fine quadrature near the origin, where the profile varies fastest
        f = np.array([O.dipole(sa, ss, 1.3, r * r)[c] for r in rs[:20001]]) * 2 * np.pi * rs[:20001]
        tail = R[2000:, c] * 2 * np.pi * rs[::10][2000:]
        num = np.trapezoid(f, rs[:20001]) + np.trapezoid(tail, rs[::10][2000:])
        assert num == pytest.approx(_closed_integral(sa, ss, 1.3, c), rel=2e-6)


def test_profile_closed_form_at_zero_and_far_away():
    sa, ss = MARBLE
    F = O.orc().orc_dipole_fdr(1.3)
    A = (1 + F) / (1 - F)
    r0 = O.dipole(sa, ss, 1.3, 0.0)
    for c in range(3):
        st = sa[c] + ss[c]
        s = math.sqrt(3 * sa[c] * st)
        zp = 1 / st
        zn = zp * (1 + 4 / 3 * A)
        # at d = 0: d+ = z+, d- = z-
        want = ss[c] / st / (4 * math.pi * s) * ((s * zp + 1) * math.exp(-s * zp) / zp ** 2
                                                 - (s * zn + 1) * math.exp(-s * zn) / zn ** 2)
        assert r0[c] == pytest.approx(want, rel=1e-13)
    assert np.all(O.dipole(sa, ss, 1.3, 100.0 ** 2) == 0.0)  # clamped beyond r*


def test_dipole_sum_is_linear_in_irradiance():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1, 1, (50, 3)).astype(np.float32)
    E = rng.random((50, 3)).astype(np.float32)
    q = rng.uniform(-1, 1, (7, 3)).astype(np.float32)
    a = O.dipole_sum(*MARBLE, 1.3, pts, E, 0.01, q)
    b = O.dipole_sum(*MARBLE, 1.3, pts, 2 * E, 0.01, q)
    np.testing.assert_allclose(b, 2 * a, rtol=1e-12)
    one = O.dipole_sum(*MARBLE, 1.3, pts[:1], E[:1], 0.01, q[:1])[0]
    d2 = float(np.sum((q[0].astype(np.float64) - pts[0]) ** 2))
    np.testing.assert_allclose(one, (1 - O.orc().orc_dipole_fdr(1.3)) / math.pi * 0.01 * O.dipole(*MARBLE, 1.3, d2) * E[0],
                               rtol=1e-12)


# ---- the octree cut (subsurface_integrator.cc:69-167) -----------------------
@pytest.mark.parametrize("name", ["sphere", "plane", "sphere_tight"])
def test_dipole_octree_matches_reference_golden(golden, name):
    """The oracle's octree cut against the reference's own Octree (construct +
    iradSubsurface + irradiance's factor), both with the reference's exp."""
    pts, E, q = golden[f"oct_{name}_pts"], golden[f"oct_{name}_E"], golden[f"oct_{name}_q"]
    prm, ref = golden[f"oct_{name}_params"], golden[f"oct_{name}_out"]
    out, vis = O.dipole_octree(prm[0:3], prm[3:6], prm[6], pts, E, prm[7], q, max_error=prm[8], fast=True)
    np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-300)
    if prm[8] >= 0.05:
        assert vis.mean() < 0.5 * len(pts)  # the cut is a cut


def _cloud(n=1500, seed=5):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    pts = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    E = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    qv = rng.normal(size=(100, 3))
    q = (qv / np.linalg.norm(qv, axis=1, keepdims=True)).astype(np.float32)
    return pts, E, q, 4 * np.pi / n


def test_dipole_octree_without_cut_is_the_exact_sum():
    """max_error = 0 accepts no interior node: the walk reaches every leaf."""
    pts, E, q, area = _cloud()
    sa, ss = MARBLE
    out, vis = O.dipole_octree(sa, ss, 1.3, pts, E, area, q, max_error=0.0)
    exact = O.dipole_sum(sa, ss, 1.3, pts, E, area, q)
    np.testing.assert_allclose(out, exact, rtol=1e-12)
    assert vis.min() >= len(pts)


def test_dipole_octree_coincident_points_share_a_leaf():
    """Points at one position (the reference would split them forever) form
    one leaf standing for all of them."""
    pts, E, q, area = _cloud(400)
    pts = np.concatenate([pts, np.repeat(pts[:3], 5, axis=0)])
    E = np.concatenate([E, np.full((15, 3), 0.7, np.float32)])
    sa, ss = MARBLE
    out, _ = O.dipole_octree(sa, ss, 1.3, pts, E, area, q, max_error=0.0)
    np.testing.assert_allclose(out, O.dipole_sum(sa, ss, 1.3, pts, E, area, q), rtol=1e-12)


def test_dipole_octree_cut_error_is_small():
    """With the reference's bound 0.05 the cut stays within 2 % of the exact
    sum on a uniformly lit sphere, and the error falls with the bound."""
    pts, E, q, area = _cloud(4000, seed=9)
    E[:] = 1.0
    sa, ss = (0.05, 0.08, 0.12), (1.0, 1.2, 1.5)
    exact = O.dipole_sum(sa, ss, 1.3, pts, E, area, q)
    e1 = np.abs(O.dipole_octree(sa, ss, 1.3, pts, E, area, q, max_error=0.05)[0] / exact - 1).max()
    e2 = np.abs(O.dipole_octree(sa, ss, 1.3, pts, E, area, q, max_error=0.005)[0] / exact - 1).max()
    assert e1 < 0.02 and e2 < e1
