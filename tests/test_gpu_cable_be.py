"""GPU parity of the passive-cell backward Euler (``bc_cable_be_run``)
against the FP64 oracle: a ragged batch of cells (every CV), the Cell front
end through ``Cell.run`` and one registered step at a time, FP32 error, and
the non-finite status of a run whose explicit axial part diverges."""

import numpy as np
import pytest

from tests.helpers import rmse
from tests.test_cable_be import one_cv_model, passive_model

pytestmark = pytest.mark.gpu


def _oracle(m, dt, n):
    from oracle.cable_be import CableBEOracle

    o = CableBEOracle(m)
    o.reset_state(m.V0)
    return o.run(dt, n)


@pytest.mark.parametrize("precision", [64, 32])
def test_ragged_batch_every_cv_matches_oracle(cuda, precision):
    import torch

    from braincell_amd.cable import CableLayout, CableRuntime

    dt, n = 0.002, 1500
    models = [passive_model(0, cv_per_branch=1), one_cv_model(), passive_model(3, cv_per_branch=1, k_leak=True),
              passive_model(1, cv_per_branch=1, cm=1.4)]
    print(f"cable backward Euler fp{precision}: {len(models)} cells, {n} steps", flush=True)
    L = CableLayout(models, probes="all")
    rt = CableRuntime(L, dt, precision=precision, threads=128)
    tr = rt.run(n)
    torch.cuda.synchronize()
    rt.check_status()
    got = tr.double().cpu().numpy()
    col = 0
    for m in models:
        ref = _oracle(m, dt, n)
        g = got[:, col:col + ref.shape[1]]
        col += ref.shape[1]
        err = np.abs(g - ref).max()
        print("cell CVs", ref.shape[1], "max |dV|", err, "RMSE", rmse(g, ref), "swing", np.ptp(ref))
        assert np.ptp(ref) > 0.5  # the clamps move the voltages
        if precision == 64:
            assert err < 1e-9
        else:
            assert rmse(g, ref) < 2e-3
    V_end = rt.V.double().cpu().numpy()
    np.testing.assert_allclose(V_end, got[-1], rtol=0, atol=0)


def _cell(precision, solver="backward_euler_mi355x"):
    from braincell_amd.cell import Cell
    from braincell_amd.filter import AllRegion, RootLocation
    from braincell_amd.mech import CableProperty, Channel, CurrentClamp, StateProbe
    from braincell_amd.policy import CVPerBranch
    from braincell_amd import workloads as W

    c = Cell(W.binary_tree_morphology(2), cv_policy=CVPerBranch(1), solver=solver, precision=precision)
    c.paint(AllRegion(), CableProperty(resting_potential=-70.0, axial_resistivity=120.0),
            Channel("IL", g_max=2e-4, E=-70.0))
    c.place(RootLocation(0.5), CurrentClamp(delay=0.2, durations=[1.0, 1.0], amplitudes=[0.4, -0.1]))
    c.place(RootLocation(0.5), StateProbe("v"))
    return c


def test_cell_run_and_update_steps_match_the_oracle(cuda):
    from braincell_amd import environ

    dt, n = 0.002, 1200
    print(f"Cell backward_euler_mi355x fp64, {n} steps", flush=True)
    c = _cell(64)
    res = c.run(dt=dt, duration=n * dt)
    m = c.lower()
    o_ref = _oracle(m, dt, n)
    soma = int(c._layout.cells[0].probes[0])
    err = np.abs(res.traces["v"][:, 0] - o_ref[:, soma]).max()
    print("Cell.run max |dV|", err, "swing", np.ptp(o_ref[:, soma]))
    assert err < 1e-9 and np.ptp(o_ref[:, soma]) > 1.0
    stepper = _cell(64)
    vs = []
    with environ.context(dt=dt):
        for i in range(100):
            with environ.context(t=i * dt):
                stepper.update()
            vs.append(float(stepper.last_trace[0, 0]))
    assert np.abs(np.asarray(vs) - res.traces["v"][:100, 0]).max() < 1e-12
    more = stepper.run(dt=dt, duration=(n - 100) * dt)
    assert np.abs(more.traces["v"][:, 0] - res.traces["v"][100:, 0]).max() < 1e-12


def test_diverging_explicit_axial_part_reports_nonfinite(cuda):
    """Four CVs on the 20 um soma make the scheme diverge at dt 0.025 (as in
    the reference): the kernel ends normally and reports NONFINITE."""
    import torch

    from braincell_amd._native import NativeError
    from braincell_amd.cable import CableLayout, CableRuntime

    print("cable backward Euler divergence, fp32, 400 steps", flush=True)
    rt = CableRuntime(CableLayout([passive_model(0)], probes="soma"), 0.025, precision=32)
    rt.run(400)
    torch.cuda.synchronize()
    with pytest.raises(NativeError):
        rt.check_status()
