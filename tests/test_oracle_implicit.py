"""Pins the oracle's linearly implicit point-neuron steps (backward_euler and
the Newton iterate of implicit_euler, ``oracle/single.py``): the reference's
own linear-decay vectors (``quad/_backward_euler_test.py:63-72``,
``quad/_implicit_test.py:65-78``), the complex-step Jacobian against finite
differences, the Jacobian's sparsity (the pattern the HIP kernel eliminates)
and the orders of convergence, 1 for theta = 1 and 2 for theta = 1/2 (the
one-stage Rosenbrock method with gamma = 1/2 is second order only with the
exact Jacobian)."""

import copy
import warnings

import numpy as np
import pytest

from braincell_amd import workloads as W
from oracle.single import NAMES, HTCPopulationOracle


def _pop(n, g_override=None, I=None, V0=None):
    pop = W.htc_population(n, seed=3)
    g = {k: v.copy() for k, v in pop["g"].items()}
    if g_override is not None:
        g = {k: np.full(n, g_override.get(k, 0.0)) for k in NAMES}
    params = {nm: {k: v for k, v in W.HTC_CHANNELS[nm].items() if k != "g_max"} for nm in NAMES}
    o = HTCPopulationOracle(g, pop["I_dc_nA"] if I is None else I, params, W.HTC_IONS)
    o.reset_state(pop["V0"] if V0 is None else V0)
    return o


def _active(n=16, steps=300):
    o = _pop(n)
    for _ in range(steps):
        o.step(0.025)
    return o


def test_linear_decay_reference_vectors():
    # dx/dt = -x / tau with tau = 10 ms, dt = 0.1 ms, x0 = 1: a leak-only
    # neuron with g / C = 0.1 / ms and E = 0
    for theta, want in ((1.0, 1.0 / 1.01), (0.5, (1.0 - 0.005) / (1.0 + 0.005))):
        o = _pop(3, g_override={"IL": 1e-4}, I=np.zeros(3), V0=np.ones(3))
        o.il_E = 0.0
        o.implicit_step(0.1, theta)
        np.testing.assert_allclose(o.V, want, rtol=0, atol=1e-13)
        if theta == 0.5:  # the reference test's bracket: implicit Euler .. exact decay
            assert np.all((o.V > 0.989) & (o.V < 0.991))


def test_complex_step_jacobian_matches_finite_differences():
    o = _active()
    y = o.state_vector()
    J = o.jacobian(y)
    for k, h in enumerate([1e-5] + [1e-7] * 9 + [1e-10]):
        yp, ym = y.copy(), y.copy()
        yp[k] += h
        ym[k] -= h
        fd = (o.rhs(yp) - o.rhs(ym)) / (2 * h)
        scale = np.abs(J[:, k]).max(axis=0, keepdims=True) + 1e-300
        assert np.max(np.abs(fd - J[:, k]) / scale) < 3e-5, k


def test_jacobian_sparsity_is_the_eliminated_pattern():
    # V couples to everything; gate k to V and itself, AHP (slot 8) to Ci
    # instead of V; Ci to V, the four Ca gates and itself
    o = _active()
    J = o.jacobian(o.state_vector())
    nz = np.abs(J).max(axis=2) > 0
    want = np.eye(11, dtype=bool)
    want[0, :] = True
    want[1:10, 0] = True
    want[8, 0], want[8, 10] = False, True
    want[10, [0, 4, 5, 6, 7]] = True
    assert np.array_equal(nz, want)


@pytest.mark.parametrize("theta,order", [(1.0, 1.0), (0.5, 2.0)])
def test_orders_of_convergence(theta, order):
    o = _active(8)
    T = 1.0

    def run(th, dt):
        q = copy.deepcopy(o)
        for _ in range(int(round(T / dt))):
            q.implicit_step(dt, th)
        return q.V

    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ref = run(0.5, 0.005 / 8)
        e = [np.abs(run(theta, dt) - ref) for dt in (0.02, 0.01, 0.005)]
    ratios = [np.median(e[i] / e[i + 1]) for i in range(2)]
    assert all(abs(np.log2(r) - order) < 0.15 for r in ratios), ratios


def test_steady_gates_and_zero_drive_give_a_fixed_point():
    # at rest with no input and no channels but the leak at its reversal, f = 0
    from oracle import channels as ch

    o = _pop(4, g_override={"IL": 1e-4}, I=np.zeros(4), V0=np.full(4, -70.0))
    o.il_E = -70.0
    o.Ci = np.full(4, float(W.HTC_IONS["ca"]["C_rest"]))
    a, b = ch.ahp_de1994(o.Ci, *o.ahp[:3])
    o.s["ahp_p"] = a / (a + b)
    y0 = o.state_vector()
    o.implicit_step(0.05, 1.0)
    np.testing.assert_allclose(o.state_vector(), y0, rtol=0, atol=1e-15)
