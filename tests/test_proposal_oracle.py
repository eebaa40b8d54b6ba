"""CPU checks that pin the oracle of the learned emission proposal.

The proposal has no reference counterpart (it replaces the cosine emission of
sources/sampler.cc:17-32), so its oracle is pinned by closed forms: the
uniform-table and one-hot-table mixtures, lambda = 1 against the plain
emission, unbiasedness of the flux factor, and finite differences of the MLP
loss.
"""
import math

import numpy as np
import pytest

import oracle as O
import proposal_ref as R

N_LIGHT = np.array([0.0, -1.0, 0.0])


def _uniform_cdf(side):
    return R.cdf_table(np.full((side * side, 1024), 1.0, np.float32))


def _basis(w):
    w = np.asarray(w, np.float64)
    u = np.cross([0.0, 1.0, 0.0], w) if abs(w[0]) > 1e-12 else np.cross([1.0, 0.0, 0.0], w)
    u /= np.linalg.norm(u)
    return u, np.cross(w, u), w


def _cell_of(d, w=N_LIGHT):
    u, v, n = _basis(w)
    ct = float(np.dot(d, n))
    phi = math.atan2(float(np.dot(d, v)), float(np.dot(d, u))) % (2 * math.pi)
    return min(int((1.0 - ct) * 32), 31), min(int(phi / (2 * math.pi) * 32), 31), ct


def test_lambda_one_is_plain_emission():
    pr = O.proposal(_uniform_cdf(4), 4, 1.0)
    for idx in range(200):
        uv, d, wt, cell = O.emission(pr, 1000, 3, 11, idx, N_LIGHT)
        uv0, d0, wt0, cell0 = O.emission(None, 1000, 3, 11, idx, N_LIGHT)
        assert wt == 1.0 and np.array_equal(d, d0) and np.array_equal(uv, uv0) and cell == cell0


def test_uniform_table_is_uniform_hemisphere():
    # lambda -> 0 with q = 1/1024: p = 1 / (2 pi), factor = (cos / pi) / p = 2 cos
    pr = O.proposal(_uniform_cdf(2), 2, 1e-12)
    for idx in range(300):
        _, d, wt, cell = O.emission(pr, 1000, 1, 5, idx, N_LIGHT)
        kz, kp, ct = _cell_of(d)
        assert abs(wt - 2.0 * ct) < 1e-6 * max(1.0, wt)
        assert abs(np.linalg.norm(d) - 1.0) < 1e-9


@pytest.mark.parametrize("k", [0, 37, 511, 1023])
def test_one_hot_table_stays_in_its_cell(k):
    prob = np.full((4, 1024), 0.0, np.float32)
    prob[:, k] = 1.0
    pr = O.proposal(R.cdf_table(prob), 2, 1e-12)
    u, v, n = _basis(N_LIGHT)
    for idx in range(200):
        _, d, wt, cell = O.emission(pr, 1000, 2, 9, idx, N_LIGHT)
        assert cell == k
        ct = float(np.dot(d, n))
        phi = math.atan2(float(np.dot(d, v)), float(np.dot(d, u))) % (2 * math.pi)
        # 1 - cos(theta) in band k // 32, phi in sector k % 32 (boundaries within rounding)
        assert k // 32 - 1e-9 <= (1.0 - ct) * 32 <= k // 32 + 1 + 1e-9
        sec = phi / (2 * math.pi) * 32
        assert k % 32 - 1e-9 <= sec <= k % 32 + 1 + 1e-9 or (k % 32 == 0 and sec > 32 - 1e-9)
        # p = q / dOmega = 1024 / (2 pi): factor = (cos / pi) * 2 pi / 1024 = cos / 512
        assert abs(wt - ct / 512.0) < 1e-9


def test_cosine_component_cells_match_directions():
    pr = O.proposal(_uniform_cdf(2), 2, 1.0 - 1e-12)
    agree = 0
    for idx in range(500):
        _, d, _, cell = O.emission(pr, 1000, 1, 3, idx, N_LIGHT)
        kz, kp, _ = _cell_of(d)
        agree += (kz * 32 + kp) == cell
    assert agree >= 498  # only exact cell boundaries may round differently


def test_flux_factor_is_unbiased():
    # E_p[(cos / pi) / p] = integral of cos / pi = 1 for any table and lambda
    rng = np.random.default_rng(4)
    prob = rng.gamma(0.3, size=(16, 1024)).astype(np.float32)
    pr = O.proposal(R.cdf_table(prob), 4, 0.25)
    w = np.array([O.emission(pr, 40000, 1, 77, i, N_LIGHT)[2] for i in range(40000)])
    assert abs(w.mean() - 1.0) < 4.0 * w.std() / math.sqrt(len(w)) + 1e-3


def test_cdf_table_properties():
    rng = np.random.default_rng(1)
    prob = rng.random((3, 1024)).astype(np.float32)
    prob[1] = 0.0
    c = R.cdf_table(prob)
    assert np.all(np.diff(c, axis=1) >= 0) and np.all(c[:, -1] == 1.0)
    np.testing.assert_allclose(np.diff(c[0], prepend=0), prob[0] / prob[0].sum(), atol=2e-6)
    np.testing.assert_allclose(np.diff(c[1], prepend=0), 1 / 1024, atol=1e-7)


def test_encoding_closed_form():
    e = R.encode([[0.0, 0.5]])[0]
    np.testing.assert_allclose(e[0:16:2], 0.0, atol=1e-15)   # sin(0)
    np.testing.assert_allclose(e[1:16:2], 1.0)               # cos(0)
    np.testing.assert_allclose(e[16], 1.0)                    # sin(pi / 2)
    np.testing.assert_allclose(e[19], -1.0, atol=1e-15)       # cos(pi)


def test_zero_output_layer_is_uniform():
    w = np.zeros(R.NPARAM)
    w[:2048] = np.random.default_rng(0).normal(size=2048)
    p = R.forward(w, np.random.default_rng(1).random((5, 2)))
    np.testing.assert_allclose(p, 1.0 / 1024, rtol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = rng.normal(scale=0.3, size=R.NPARAM)
    rec = np.column_stack([rng.random(24), rng.random(24), rng.integers(0, 1024, 24), rng.random(24) + 0.1])
    loss, g = R.loss_and_grad(w, rec, emulate=False)
    assert abs(loss - R.loss_only(w, rec)) < 1e-12
    idx = np.concatenate([rng.choice(np.arange(0, 2048), 6), rng.choice(np.arange(2048, 2112), 3),
                          rng.choice(np.arange(2112, 10432), 8), rng.choice(np.arange(10432, 75968), 8),
                          rng.choice(np.arange(75968, R.NPARAM), 4)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (R.loss_only(wp, rec) - R.loss_only(wm, rec)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2.0, 1e-3], np.float32)
    w, m, v = R.adam(np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32), g, 1, 0.01)
    np.testing.assert_allclose(w, -0.01 * np.sign(g), rtol=1e-4)
