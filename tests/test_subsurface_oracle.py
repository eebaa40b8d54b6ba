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
