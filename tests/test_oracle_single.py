"""Pins the point-neuron oracle (ind_exp_euler on SingleCompartment)."""

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


def test_dv_dt_equals_I_over_C():
    # _single_compartment/base_test.py:343-360: no channels, I = C * 1 mV/ms
    area = 2 * np.pi * 5.0 * 10.0 * 1e-8
    o = _pop(3, g_override={}, I=np.full(3, 1.0 * area * 1e3), V0=np.full(3, -65.0))
    o.step(0.1)
    np.testing.assert_allclose(o.V, -65.0 + 0.1, rtol=0, atol=1e-12)


def test_leak_matches_analytic_solution():
    # base_test.py:380-425: exp-Euler is exact for the linear leak equation
    g_l, E = 0.0003, -70.0
    area = 2 * np.pi * 5.0 * 10.0 * 1e-8
    I_nA = np.asarray([0.0, 0.01, 0.02])
    o = _pop(3, g_override={"IL": g_l}, I=I_nA, V0=np.full(3, -60.0))
    gm = g_l * 1e3
    Iu = I_nA * 1e-3 / area
    Vinf = E + Iu / gm
    for n in range(1, 401):
        o.step(0.025)
        np.testing.assert_allclose(o.V, Vinf + (-60.0 - Vinf) * np.exp(-gm * n * 0.025), rtol=0, atol=1e-9)


def test_self_jacobians_match_finite_differences():
    o = _pop(16)
    for _ in range(300):
        o.step(0.025)
    d = o.derivatives(o.V, o.s, o.Ci)
    h = 1e-6
    fV = lambda v: o.derivatives(v, o.s, o.Ci)["V"][0]
    np.testing.assert_allclose((fV(o.V + h) - fV(o.V - h)) / (2 * h), d["V"][1], rtol=1e-6, atol=1e-9)
    for k in o.s:
        sp = dict(o.s, **{k: o.s[k] + 1e-7})
        sm = dict(o.s, **{k: o.s[k] - 1e-7})
        fd = (o.derivatives(o.V, sp, o.Ci)[k][0] - o.derivatives(o.V, sm, o.Ci)[k][0]) / 2e-7
        np.testing.assert_allclose(fd, d[k][1], rtol=1e-6, atol=1e-9)
    hc = 1e-10
    fd = (o.derivatives(o.V, o.s, o.Ci + hc)["Ci"][0] - o.derivatives(o.V, o.s, o.Ci - hc)["Ci"][0]) / (2 * hc)
    np.testing.assert_allclose(fd, d["Ci"][1], rtol=1e-5)


def test_steady_state_gates_are_fixed_points_of_their_update():
    o = _pop(8)
    d = o.derivatives(o.V, o.s, o.Ci)
    for k in ("na_p", "na_q", "kd_p", "cl_p", "cl_q", "ct_p", "ct_q", "hcn_p", "ahp_p"):
        assert np.max(np.abs(d[k][0])) < 1e-12, k


def test_population_fires():
    o = _pop(32)
    tr = o.run(0.025, 4000)
    assert tr.max() > 20.0
