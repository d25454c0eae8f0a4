"""Non-finite values injected into a host-prepared state end a run with
status NONFINITE while every kernel exits normally, for the staggered step
(FP64 and FP32, NaN in V and in a gate) and for point neurons; a reset
state then runs clean again."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from braincell_amd._native import NativeError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("where", ["V", "gate"])
def test_nan_in_state_raises_nonfinite(cuda, precision, where):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    models = [W.uneven_tree_cell(i, cmin=200, cmax=400) for i in range(3)]
    L = StaggerLayout(models, probes="soma")
    rt = StaggerRuntime(L, 0.025, precision=precision)
    rt.enable_spikes(0.0, cap=8)
    print(f"nan injection: precision {precision}, {where}", flush=True)
    # NaN in the second cell's state, prepared on the host side of the run
    cv = int(L.cell_cv0[1]) + 5
    if where == "V":
        rt.V[cv] = float("nan")
    else:
        rt.S[0 * L.n_cv_total + cv] = float("nan")  # Na activation gate
    tr = rt.run(40, chunk=20)
    torch.cuda.synchronize()
    assert int(rt.status.item()) & 1
    with pytest.raises(NativeError) as e:
        rt.check_status()
    assert e.value.code == 4
    # the other cells were integrated (finite soma traces), the kernel ended normally
    v = tr.double().cpu().numpy()
    assert np.isfinite(v[:, 0]).all() and np.isfinite(v[:, 2]).all()
    # a reset state runs clean again
    rt.reset_state()
    rt.run(40, chunk=20)
    torch.cuda.synchronize()
    rt.check_status()


def test_nan_in_point_population_raises_nonfinite(cuda):
    import torch
    from braincell_amd.single import SingleCompartment

    pop = W.htc_population(300, seed=3)
    channels = {nm: dict(W.HTC_CHANNELS[nm], g_max=pop["g"][nm]) for nm in pop["names"]}
    sc = SingleCompartment(300, channels=channels, E_Na=W.HTC_IONS["E_Na"], E_K=W.HTC_IONS["E_K"],
                           ca=W.HTC_IONS["ca"], V0=pop["V0"], precision=32)
    sc.init_state()
    sc.set_input(pop["I_dc_nA"])
    sc._dev["V"][17] = float("nan")
    with pytest.raises(NativeError) as e:
        sc.run(dt=0.025, duration=2.0)
    assert e.value.code == 4
    sc.reset_state(pop["V0"])
    sc.run(dt=0.025, duration=2.0)
    torch.cuda.synchronize()


def test_undersized_synapse_plan_is_reported(cuda):
    """A synapse descriptor whose LDS plan is smaller than a pack's
    synapses (forced from the host): the kernel skips those synapses, ends
    normally and reports status 4."""
    import torch
    from tests.test_gpu_network import _markov_net

    net = _markov_net(32)
    tabs = net.build(0.025, chunk=200)
    assert tabs.syn.max_syn > 1
    print(f"undersized synapse plan: max_syn {tabs.syn.max_syn} -> 1", flush=True)
    rt = tabs.runtime(80)
    rt.synapse_desc.max_syn = rt.synapse_desc.max_syn_rows = 1
    rt.run(80, chunk=tabs.launch_steps)
    torch.cuda.synchronize()
    assert int(rt.status.item()) & 4
    with pytest.raises(NativeError, match="synapse"):
        rt.check_status()
