"""Spike output of plain runs (get_spike, braincell/_base.py:421-431; set
every step by Cell._update_dynamics, _multi_compartment/cell.py:2384-2386)
against the oracle's threshold crossings, in FP64 and FP32, through the
runtime, hipGraph replays and the Cell / CellBatch front ends."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import run_oracle, soma_cv

pytestmark = pytest.mark.gpu

CELLS = (30, 49, 14, 27)  # 388, 276, 468, 204 CVs; the first three spike under their clamps
STEPS = 2400


def crossings(v0, trace, th=0.0):
    """Steps n with V_n < th <= V_{n+1} (V_0 = v0, V_{n+1} = trace[n])."""
    prev = np.concatenate([[v0], trace[:-1]])
    return np.nonzero((prev < th) & (trace >= th))[0]


@pytest.fixture(scope="module")
def oracle_steps():
    out = []
    for i in CELLS:
        m = W.uneven_tree_cell(i)
        ref, _ = run_oracle(m, STEPS, probe_cvs=[soma_cv(m)])
        out.append(crossings(m.V0, ref[:, 0]))
    assert sum(len(s) > 0 for s in out) == 3
    return out


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("mode", ["one-launch", "chunks-of-7", "graph"])
def test_spike_steps_match_oracle(cuda, oracle_steps, precision, mode):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    models = [W.uneven_tree_cell(i) for i in CELLS]
    L = StaggerLayout(models, probes="soma")
    rt = StaggerRuntime(L, 0.025, precision=precision)
    rt.enable_spikes(0.0, cap=64)
    print(f"spikes: precision {precision}, {mode}", flush=True)
    if mode == "graph":
        tr = rt.new_trace(STEPS)
        g = rt.make_graph(STEPS // 200, 200, tr)
        g.launch(rt.stream_handle())
        rt.step_index += STEPS
    else:
        tr = rt.run(STEPS, chunk=STEPS if mode == "one-launch" else 7)
    torch.cuda.synchronize()
    rt.check_status()
    got = rt.spike_steps()
    v = tr.double().cpu().numpy()
    for k, i in enumerate(CELLS):
        # the record is the kernel trace's own crossings ...
        np.testing.assert_array_equal(got[k], crossings(models[k].V0, v[:, k]))
        # ... and the oracle's
        np.testing.assert_array_equal(got[k], oracle_steps[k]), (i, got[k], oracle_steps[k])


def test_spike_record_restarts_and_counts_past_cap(cuda):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    m = W.uneven_tree_cell(30)
    L = StaggerLayout([m], probes="soma")
    rt = StaggerRuntime(L, 0.025, precision=32)
    rt.enable_spikes(0.0, cap=2)
    rt.run(STEPS, chunk=200)
    torch.cuda.synchronize()
    assert int(rt._spk_buf["count"][0].item()) > 2
    with pytest.raises(RuntimeError, match="exceed"):
        rt.spike_steps()
    full = rt._spk_buf["steps"][0].cpu().numpy()
    rt.reset_state()
    assert int(rt._spk_buf["count"][0].item()) == 0
    rt.run(STEPS, chunk=200)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(rt._spk_buf["steps"][0].cpu().numpy(), full)


def test_cell_and_batch_front_ends_report_spike_times(cuda):
    from braincell_amd.cell import Cell, CellBatch
    from braincell_amd.filter import AllRegion, RootLocation
    from braincell_amd.mech import CableProperty, Channel, CurrentClamp
    from braincell_amd.morph import Branch, Morphology
    from braincell_amd.policy import CVPerBranch
    from oracle.staggered import StaggeredOracle

    mo = Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma"))
    mo.attach(0, Branch.from_lengths(lengths=[150.0], radii=[1.0, 0.6], type="dendrite"))
    c = Cell(mo, cv_policy=CVPerBranch(4), precision=64)
    c.paint(AllRegion(), CableProperty(resting_potential=-70.0), Channel("IL", g_max=3e-4, E=-70.0),
            Channel("Na_Ba2002", g_max=0.12), Channel("KDR_Ba2002", g_max=0.036))
    c.place(RootLocation(0.5), CurrentClamp(delay=5.0, durations=[30.0], amplitudes=[0.3]))
    res = c.run(dt=0.025, duration=50.0)
    m = c.lower()
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    ref = o.run(0.025, res.time.size, probe_cvs=[soma_cv(m)])[:, 0]
    want = res.time[crossings(m.V0, ref)]
    assert want.size >= 2
    np.testing.assert_array_equal(res.spikes["V"][0], want)
    # a second run continues from the cell's state; spike times are absolute
    res2 = c.run(dt=0.025, duration=10.0)
    assert res2.time[0] == pytest.approx(50.0)
    assert all(t >= 50.0 for t in res2.spikes["V"][0])
    assert c.spike.shape == (1,)
    # CellBatch: the same cell twice plus a resting copy without the clamp
    rest = c.lower()
    rest.clamps = []
    b = CellBatch([m, m, rest], precision=64)
    rb = b.run(dt=0.025, duration=50.0)
    np.testing.assert_array_equal(rb.spikes["V_soma"][0], want)
    np.testing.assert_array_equal(rb.spikes["V_soma"][1], want)
    assert rb.spikes["V_soma"][2].size == 0
