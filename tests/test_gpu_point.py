"""GPU parity of the point-neuron (ind_exp_euler) kernel against the oracle."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import rmse, spike_times

pytestmark = pytest.mark.gpu


def _htc(n, seed=1):
    pop = W.htc_population(n, seed=seed)
    channels = {nm: dict(W.HTC_CHANNELS[nm], g_max=pop["g"][nm]) for nm in pop["names"]}
    return pop, channels


def _run(n, steps, precision, n_trace):
    import torch
    from braincell_amd.single import SingleCompartment

    pop, channels = _htc(n)
    sc = SingleCompartment(n, channels=channels, E_Na=W.HTC_IONS["E_Na"], E_K=W.HTC_IONS["E_K"],
                           ca=W.HTC_IONS["ca"], V0=pop["V0"], precision=precision)
    sc.init_state()
    sc.set_input(pop["I_dc_nA"])
    tr = sc.run(dt=0.025, duration=steps * 0.025, n_trace=n_trace)
    torch.cuda.synchronize()
    return tr.double().cpu().numpy(), pop, channels, sc


def _oracle(pop, channels, n, steps):
    from oracle.single import HTCPopulationOracle

    g = {nm: pop["g"][nm][:n] for nm in pop["names"]}
    params = {nm: {k: v for k, v in W.HTC_CHANNELS[nm].items() if k != "g_max"} for nm in pop["names"]}
    o = HTCPopulationOracle(g, pop["I_dc_nA"][:n], params, W.HTC_IONS)
    o.reset_state(pop["V0"][:n])
    return o.run(0.025, steps)


def test_point_fp64_matches_oracle(cuda):
    n, steps = 64, 4000
    got, pop, ch, _ = _run(n, steps, 64, n)
    ref = _oracle(pop, ch, n, steps)
    err = np.abs(got - ref).max()
    print("point fp64 max |dV|:", err)
    assert err < 1e-6


def test_point_fp32_rmse_and_spike_times(cuda):
    n, steps = 256, 4000
    got, pop, ch, sc = _run(n, steps, 32, n)
    ref = _oracle(pop, ch, n, steps)
    worst = max(rmse(got[:, i], ref[:, i]) for i in range(n))
    print("point fp32 worst RMSE over 100 ms:", worst)
    assert worst <= 0.05
    nsp = 0
    for i in range(n):
        a, b = spike_times(got[:, i], 0.025), spike_times(ref[:, i], 0.025)
        assert len(a) == len(b)
        if len(b):
            assert np.max(np.abs(a - b)) <= 0.025 + 1e-9
        nsp += len(b)
    print("spikes:", nsp)
    assert nsp > 0


def test_point_fp32_long_run_within_fp32_noise_floor(cuda):
    """Over 200 ms a few of these fast-firing neurons are chaotic: the FP64
    oracle with FP32-sized relative noise injected every step diverges on
    them too.  The FP32 kernel may fail the 0.05 mV bar only on as many
    neurons as that noise floor does."""
    from oracle.single import HTCPopulationOracle

    n, steps = 256, 8000
    got, pop, ch, _ = _run(n, steps, 32, n)
    ref = _oracle(pop, ch, n, steps)
    fail_gpu = sum(rmse(got[:, i], ref[:, i]) > 0.05 for i in range(n))
    params = {nm: {k: v for k, v in W.HTC_CHANNELS[nm].items() if k != "g_max"} for nm in pop["names"]}
    o = HTCPopulationOracle(pop["g"], pop["I_dc_nA"], params, W.HTC_IONS)
    o.reset_state(pop["V0"])
    rng = np.random.default_rng(5)
    noisy = np.empty((steps, n))
    for k in range(steps):
        o.step(0.025)
        o.V = o.V * (1 + 6e-8 * rng.standard_normal(n))
        for key in o.s:
            o.s[key] = o.s[key] * (1 + 6e-8 * rng.standard_normal(n))
        noisy[k] = o.V
    fail_floor = sum(rmse(noisy[:, i], ref[:, i]) > 0.05 for i in range(n))
    print(f"200 ms: fp32 kernel fails {fail_gpu}/{n}, FP64 + FP32-noise floor fails {fail_floor}/{n}")
    assert fail_gpu <= fail_floor + 2
