"""Pins the FP64 numpy oracle of the learned Newton guess (CPU only)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402

import oracle_lib as O  # noqa: E402


def _params(seed=3, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, M.NUM_PARAMS)
    W1 = p[M.OFF_W1:M.OFF_B1].reshape(64, 8)
    W1[:, 6:] = 0.0
    return p


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(5)
    p = _params()
    x = rng.normal(0.0, 1.0, (6, 24))
    t = rng.normal(0.0, 0.5, 24)
    flag = np.zeros(24, dtype=np.int32)
    flag[[3, 10, 17]] = [1, 2, 1]  # rejected samples take no part
    l0, g = M.loss_and_grad(p, x, t, flag)
    assert l0 == pytest.approx(M.loss(p, x, t, flag), rel=1e-14)
    idx = np.concatenate([rng.choice(M.NUM_PARAMS, 60, replace=False),
                          [M.OFF_B1, M.OFF_W2 + 5 * 64 + 9, M.OFF_B2 + 63, M.OFF_W3, M.OFF_B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss(pp, x, t, flag) - M.loss(pm, x, t, flag)) / (2 * h)
        assert abs(g[k] - fd) <= 1e-7 + 1e-6 * abs(fd), (k, g[k], fd)
    # padded input columns 6, 7 of W1 never receive gradient
    assert np.all(g[M.OFF_W1:M.OFF_B1].reshape(64, 8)[:, 6:] == 0.0)


def test_huber_gradient_matches_central_differences():
    rng = np.random.default_rng(7)
    p = _params(seed=8)
    x = rng.normal(0.0, 1.0, (6, 20))
    # targets around the outputs: errors on both sides of the threshold
    t = M.forward(p, x)[0] + rng.normal(0.0, 0.4, 20)
    flag = np.zeros(20, dtype=np.int32)
    h = 0.25
    l0, g = M.loss_and_grad(p, x, t, flag, huber=h)
    assert l0 == pytest.approx(M.loss(p, x, t, flag, huber=h), rel=1e-14)
    y, _, _ = M.forward(p, x)
    assert np.any(np.abs(y - t) > h) and np.any(np.abs(y - t) < h)
    step = 1e-6
    for k in rng.choice(M.NUM_PARAMS, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += step
        pm[k] -= step
        fd = (M.loss(pp, x, t, flag, huber=h) - M.loss(pm, x, t, flag, huber=h)) / (2 * step)
        assert abs(g[k] - fd) <= 1e-7 + 1e-6 * abs(fd), (k, g[k], fd)


def test_rejected_samples_do_not_change_loss():
    rng = np.random.default_rng(6)
    p = _params()
    x = rng.normal(0.0, 1.0, (6, 10))
    t = rng.normal(0.0, 0.5, 10)
    flag = np.array([0, 1, 0, 2, 0, 0, 1, 0, 0, 0], dtype=np.int32)
    t2 = t.copy()
    t2[flag != 0] = 1e6
    assert M.loss(p, x, t, flag) == M.loss(p, x, t2, flag)
    ok = flag == 0
    assert M.loss(p, x, t, flag) == pytest.approx(M.loss(p, x[:, ok], t[ok], flag[ok]), rel=1e-14)


def test_forward_one_unit_network():
    """With a single active path the network is d relu(c relu(a x0 + b)) + e
    (one sample on each side of the first unit's kink)."""
    p = np.zeros(M.NUM_PARAMS)
    a, b, c, d, e = 0.7, -0.2, 1.3, -2.0, 0.25
    p[M.OFF_W1 + 0 * 8 + 0] = a
    p[M.OFF_B1 + 0] = b
    p[M.OFF_W2 + 5 * 64 + 0] = c   # hidden-2 unit 5 reads hidden-1 unit 0
    p[M.OFF_W3 + 5] = d
    p[M.OFF_B3] = e
    x = np.array([[0.5, -1.0], [9, 9], [9, 9], [9, 9], [9, 9], [9, 9]], dtype=np.float64)
    y, h1, h2 = M.forward(p, x)
    relu = lambda v: np.maximum(v, 0.0)
    np.testing.assert_allclose(y, d * relu(c * relu(a * x[0] + b)) + e, rtol=1e-15)
    assert y[1] == e and y[0] != e
    assert np.all(h2[:5] == 0.0) and np.all(h2[6:] == 0.0)


def test_splitmix_reference_sequence():
    """SplitMix64 from state 0 (Vigna's reference splitmix64.c)."""
    s, out = 0, []
    for _ in range(3):
        out.append(M.splitmix(s))
        s = (s + 0x9E3779B97F4A7C15) & M.MASK64
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_sample_ranges_and_labels():
    prim = M.sample_prim(11, range(400))
    for k in (0, 2, 3, 5):
        assert np.all((prim[k] >= 1e-3 * (1 - 1e-12)) & (prim[k] <= 1e3 * (1 + 1e-12)))
    assert np.all(np.abs(prim[[1, 4]]) <= 5.0)
    # log-uniform: log10 rho spreads over (-3, 3) with mean near 0
    assert abs(np.mean(np.log10(prim[0]))) < 0.3
    left, right = M.complete_pairs(prim)
    vac = M.vacuum(prim)
    pM, uM, it, conv = O.exact_middle_state(left[:, ~vac], right[:, ~vac], 2, 0)
    assert np.all(conv == 1) and np.all(pM > 0)
    # equal left and right states: the star pressure is their pressure
    same = prim.copy()
    same[3:] = same[:3]
    l2, r2 = M.complete_pairs(same)
    p2, _, _, _ = O.exact_middle_state(l2, r2, 2, 0)
    np.testing.assert_allclose(p2, same[2], rtol=1e-7)
