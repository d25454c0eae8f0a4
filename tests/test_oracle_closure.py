"""Pins the closure oracle: MLP gradient against central differences, the
plain gate update against the channel closed forms."""

import numpy as np

from oracle import channels as ch
from oracle import closure as oc


def _problem(seed=0, n=16):
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(5061) * 0.3
    Xn = rng.standard_normal((n, 8))
    Dn = rng.standard_normal((n, 5))
    Yn = rng.standard_normal((n, 5))
    return P, Xn, Dn, Yn


def test_mlp_gradient_matches_central_differences():
    P, Xn, Dn, Yn = _problem()
    _, g = oc.loss_and_grad(P, Xn, Dn, Yn)
    rng = np.random.default_rng(1)
    # every block of the parameter vector: W1, b1, W2, b2, W3, b3
    idx = np.concatenate([rng.choice(np.arange(a, b), min(6, b - a), replace=False)
                          for a, b in ((0, 512), (512, 576), (576, 4672), (4672, 4736), (4736, 5056), (5056, 5061))])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (oc.loss_and_grad(Pp, Xn, Dn, Yn)[0] - oc.loss_and_grad(Pm, Xn, Dn, Yn)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_unpack_covers_every_parameter_once():
    P = np.arange(5061, dtype=np.float64)
    flat = np.concatenate([np.concatenate([W.ravel(), b]) for W, b in oc.unpack(P)])
    assert np.array_equal(flat, P)


def test_plain_gates_are_the_exponential_relaxation():
    rng = np.random.default_rng(2)
    V = rng.uniform(-90, 40, 50)
    x = rng.uniform(0, 1, (5, 50))
    dt = 0.1
    got = oc.plain_gates(V, x, dt)
    ap, bp, aq, bq = ch.na_ba2002(V, -50.0)
    s = ap + bp
    ref = ap / s + (x[0] - ap / s) * np.exp(-dt * s)
    assert np.allclose(got[0], ref, rtol=1e-12, atol=1e-14)
    pi, pt, qi, qt = ch.cal_is2008(V, 0.0)
    assert np.allclose(got[4], qi + (x[4] - qi) * np.exp(-dt / qt), rtol=1e-12, atol=1e-14)


def test_features_layout():
    f = oc.features(np.array([-60.0, -50.0]), np.zeros((5, 2)) + 0.5, np.array([1e-4, 2e-4]), 0.1)
    assert f.shape == (2, 8)
    assert np.allclose(f[:, 0], [-60, -50]) and np.allclose(f[:, 6], np.log([1e-4, 2e-4])) and np.all(f[:, 7] == 0.1)
