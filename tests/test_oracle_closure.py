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


def test_closure_judge_spike_timing_and_rmse_helpers():
    """The judge's helpers on hand-made traces: a spike at step i of a trace
    with step dt is at (i + 1) dt; spike errors compare counts first."""
    from braincell_amd.closure import soma_rmse, spike_error, spike_times_ms

    v = np.full(10, -70.0)
    v[3] = 10.0  # V_2 < 0 <= V_3: step 3 -> (3 + 1) dt
    assert np.allclose(spike_times_ms(v, 0.1), [0.4])
    fine = np.full((200, 3), -70.0)
    fine[79:, 0] = 20.0  # crossing at row 79 -> 0.4 ms at dt 0.005
    fine[59:, 1] = 20.0  # 0.3 ms
    coarse = np.full((10, 3), -70.0)
    coarse[3:, 0] = 20.0  # 0.4 ms: exact
    coarse[4:, 1] = 20.0  # 0.5 ms: 0.2 ms late
    coarse[2:, 2] = 20.0  # a spike the fine trace does not have
    cnt, err = spike_error(coarse, fine, 0.1, 0.005)
    assert cnt.tolist() == [1, 1, 0]
    assert np.allclose(err[:2], [0.0, 0.2]) and np.isinf(err[2])
    rm = soma_rmse(coarse, fine, 20)
    ref = np.sqrt(np.mean((coarse - fine[19::20]) ** 2, axis=0))
    assert np.allclose(rm, ref)
    assert rm[0] == 0.0
