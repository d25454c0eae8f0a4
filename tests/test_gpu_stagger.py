"""GPU parity of the staggered HIP kernel against the FP64 oracle."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import rmse, run_gpu, run_oracle, soma_cv, spike_times

pytestmark = pytest.mark.gpu


def test_single_tree_fp64_all_cvs_match_oracle(cuda):
    m = W.single_tree_cell(0)
    steps = 4000
    ref, _ = run_oracle(m, steps)
    got, _ = run_gpu([m], steps, precision=64, probes="all")
    err = np.abs(got - ref).max()
    print("fp64 max |dV| over 4000 steps x 256 CVs:", err)
    assert err < 1e-8


def test_backsub_sweep_fp32_all_cvs_match_oracle(cuda):
    # the sequential back substitution of the chain tree (BC_STEP_BACKSUB_SWEEP,
    # compiled for the FP32 kernels) on every CV of the c0 tree
    m = W.single_tree_cell(0)
    steps = 4000
    ref, _ = run_oracle(m, steps)
    got, rt = run_gpu([m], steps, precision=32, probes="all", backsub="sweep")
    assert rt.L.step_flags == 8
    e = np.sqrt(np.mean((got - ref) ** 2, axis=0))
    print("fp32 sweep back substitution, worst per-CV RMSE over 4000 steps x 256 CVs:", e.max())
    assert e.max() <= 0.05


def test_backsub_sweep_packed_batch_matches_jumping(cuda):
    # several cells per pack (root chains on different levels): the two
    # variants against each other, the sweep against the oracle; FP64 has no
    # sweep kernel
    from braincell_amd._native import NativeError

    models = [W.uneven_tree_cell(i, cmin=200, cmax=700) for i in range(6)]
    steps, dt = 1200, 0.025
    a, _ = run_gpu(models, steps, precision=32, probes="soma", rows_per_pack=2600, backsub="jump")
    b, rt = run_gpu(models, steps, precision=32, probes="soma", rows_per_pack=2600, backsub="sweep")
    assert rt.L.n_packs < len(models)
    # rounding differences of the two summation orders grow through spike
    # upstrokes; the bound sits far below the 0.05 mV gate
    d = np.abs(a - b).max()
    print(f"fp32: sweep vs pointer jumping, max |dV| {d:.2e} mV")
    assert d < 2e-2
    for i, m in enumerate(models):
        assert len(spike_times(a[:, i], dt)) == len(spike_times(b[:, i], dt))
    for i, m in enumerate(models[:2]):
        ref, _ = run_oracle(m, steps, probe_cvs=[soma_cv(m)])
        assert rmse(b[:, i], ref[:, 0]) <= 0.05
    with pytest.raises(NativeError, match="sequential back substitution"):
        run_gpu(models[:1], 10, precision=64, probes="soma", backsub="sweep")


def test_single_tree_fp32_soma_rmse(cuda):
    m = W.single_tree_cell(0)
    steps = 4000
    cv = soma_cv(m)
    ref, _ = run_oracle(m, steps, probe_cvs=[cv])
    got, _ = run_gpu([m], steps, precision=32, probes="soma")
    e = rmse(got[:, 0], ref[:, 0])
    print("fp32 soma RMSE (mV):", e)
    assert e <= 0.05


def test_uneven_batch_fp32_soma_rmse_and_spikes(cuda):
    models = [W.uneven_tree_cell(i, cmin=200, cmax=800) for i in range(4)]
    steps, dt = 4000, 0.025
    got, _ = run_gpu(models, steps, precision=32, probes="soma", rows_per_pack=2600)
    n_spk = 0
    for i, m in enumerate(models):
        ref, _ = run_oracle(m, steps, probe_cvs=[soma_cv(m)])
        e = rmse(got[:, i], ref[:, 0])
        st_g, st_r = spike_times(got[:, i], dt), spike_times(ref[:, 0], dt)
        print(f"cell {i}: RMSE {e:.2e} mV, spikes gpu {len(st_g)} oracle {len(st_r)}")
        assert e <= 0.05
        assert len(st_g) == len(st_r)
        if len(st_r):
            assert np.max(np.abs(st_g - st_r)) <= dt + 1e-9
        n_spk += len(st_r)
    assert n_spk > 0


def test_uneven_batch_fp64_matches_oracle(cuda):
    # FP64 rounding differences (solve order, exp/log implementations) are
    # amplified by spike upstrokes (~100 mV/ms); 1e-5 mV bounds them with a
    # margin and is four orders below the FP32 gate.
    models = [W.uneven_tree_cell(i, cmin=200, cmax=600) for i in range(3)]
    steps = 1200
    got, _ = run_gpu(models, steps, precision=64, probes="soma", rows_per_pack=2600)
    for i, m in enumerate(models):
        ref, _ = run_oracle(m, steps, probe_cvs=[soma_cv(m)])
        err = np.abs(got[:, i] - ref[:, 0]).max()
        print(f"cell {i}: fp64 max |dV| {err:.2e} mV")
        assert err < 1e-5


def test_cell_alone_equals_cell_in_pack_bitwise(cuda):
    models = [W.uneven_tree_cell(i, cmin=200, cmax=500) for i in range(3)]
    together, _ = run_gpu(models, 400, precision=32, probes="soma", rows_per_pack=4000)
    for i, m in enumerate(models):
        alone, _ = run_gpu([m], 400, precision=32, probes="soma")
        assert np.array_equal(alone[:, 0], together[:, i])


def test_chunked_and_graph_runs_equal_one_launch(cuda):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    models = [W.uneven_tree_cell(i, cmin=200, cmax=500) for i in range(5)]
    one, _ = run_gpu(models, 600, precision=32, rows_per_pack=1500)
    chunked, _ = run_gpu(models, 600, precision=32, rows_per_pack=1500, chunk=100)
    assert np.array_equal(one, chunked)
    L = StaggerLayout(models, rows_per_pack=1500)
    rt = StaggerRuntime(L, 0.025, precision=32)
    trace = rt.new_trace(600)
    g = rt.make_graph(6, 100, trace)
    for rep in range(3):
        rt.reset_state()
        torch.cuda.synchronize()
        g.launch(rt.stream_handle())
        torch.cuda.synchronize()
        assert np.array_equal(trace.double().cpu().numpy(), one), f"replay {rep}"


def _ahp_cell():
    """Config 0 plus CaL, AHP and a CalciumDetailed pool on the soma, clamped
    hard enough to spike (Ca influx drives the AHP gate)."""
    import dataclasses

    m = W.single_tree_cell(0)
    nb = m.n_branches
    soma = np.asarray([b.type == "soma" for b in m.morph.branches])
    ch = dict(m.channels)
    ch["CaL_IS2008"] = {"g_branch": np.where(soma, 0.01, 0.0), "params": {}}
    ch["AHP_De1994"] = {"g_branch": np.where(soma, 0.02, 0.0), "params": {}}
    return dataclasses.replace(m, channels=ch, ca=W._ca_pool(soma), clamps=[(0, 0.5, 2.0, 60.0, 0.6)]).validate()


@pytest.mark.parametrize("order", ["family", "integration"])
def test_update_orders_match_oracle_fp64(cuda, order):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime
    from oracle.staggered import StaggeredOracle

    m = _ahp_cell()
    n = 1600
    L = StaggerLayout([m], probes="soma")
    rt = StaggerRuntime(L, 0.025, precision=64)
    rt.flags = 2 if order == "integration" else 0
    got = rt.run(n, chunk=400).double().cpu().numpy()[:, 0]
    torch.cuda.synchronize()
    rt.check_status()
    o = StaggeredOracle(m, update_order=order)
    o.reset_state(m.V0)
    ref = o.run(0.025, n, probe_cvs=[soma_cv(m)])[:, 0]
    assert len(spike_times(ref, 0.025)) >= 2
    assert np.max(np.abs(got - ref)) < 1e-6
    # the order matters on this cell: the other order's oracle is far away
    other = StaggeredOracle(m, update_order="family" if order == "integration" else "integration")
    other.reset_state(m.V0)
    alt = other.run(0.025, n, probe_cvs=[soma_cv(m)])[:, 0]
    assert np.max(np.abs(alt - ref)) > 1e-3


@pytest.mark.parametrize("n_cv", [1, 3, 4])
def test_interior_attachment_fp64_matches_oracle(cuda, n_cv):
    """Children attached inside their parent branch (0 < parent_x < 1),
    at x = 0 and with a reversed walk: all CVs against the oracle."""
    from tests.helpers import interior_cell

    models = [interior_cell(s, n_cv) for s in range(3)]
    steps = 1200
    got, _ = run_gpu(models, steps, precision=64, probes="all")
    col = 0
    for m in models:
        ref, _ = run_oracle(m, steps)
        n = ref.shape[1]
        err = np.abs(got[:, col:col + n] - ref).max()
        print(f"interior cell ({n} CVs): fp64 max |dV| {err:.2e} mV, V range {ref.min():.1f}..{ref.max():.1f}")
        assert err < 1e-5  # FP64 rounding amplified by spike upstrokes, as above
        col += n
    assert ref.max() > 0.0  # the clamped soma spikes


def _n144_cell():
    from braincell_amd.model import CellModel
    from braincell_amd.swc import read_swc

    from braincell_amd.workloads import N144_SWC

    morph = read_swc(N144_SWC)
    nb = morph.n_branches
    return CellModel(
        morph=morph, cv_per_branch=4, cm=1.0, ra=1000.0, V0=-65.0,
        channels={"Na_Ba2002": {"g_branch": np.full(nb, 0.12), "params": {}},
                  "KDR_Ba2002": {"g_branch": np.full(nb, 0.036), "params": {}},
                  "IL": {"g_branch": np.full(nb, 0.0003), "params": {"E": -54.3}}},
        clamps=[(0, 0.85, 3.0, 2.0, 0.7)],
    ).validate()


def test_swc_cell_matches_oracle(cuda):
    """The reference benchmark's 868-CV morphology (n144.swc, 4 CVs per
    branch, children at the soma's midpoint) with Na/KDR/leak: FP64 all CVs
    and FP32 soma against the oracle over 20 ms."""
    m = _n144_cell()
    steps = 800
    ref, _ = run_oracle(m, steps)
    assert ref.shape[1] == 868
    got64, _ = run_gpu([m], steps, precision=64, probes="all")
    err = np.abs(got64 - ref).max()
    cv = soma_cv(m)
    got32, _ = run_gpu([m], steps, precision=32, probes="soma")
    e32 = rmse(got32[:, 0], ref[:, cv])
    print(f"n144: fp64 max |dV| {err:.2e} mV, fp32 soma RMSE {e32:.2e} mV, soma peak {ref[:, cv].max():.1f} mV")
    assert err < 1e-5
    assert e32 <= 0.05


def test_n144_hh_benchmark_cell_matches_oracle(cuda):
    """The reference's simulator-comparison cell (n144.swc, 868 CVs,
    Hodgkin-Huxley 1952 at 6.3 degC): FP64 all CVs, and an FP32 batch of
    three clamp amplitudes at the soma, against the oracle over 20 ms."""
    from braincell_amd import workloads as W

    steps = 800
    m = W.n144_hh_cell(0.7)
    ref, _ = run_oracle(m, steps)
    got64, _ = run_gpu([m], steps, precision=64, probes="all")
    err = np.abs(got64 - ref).max()
    print(f"n144 HH: fp64 max |dV| {err:.2e} mV")
    assert err < 1e-5
    amps = W.n144_amplitudes(3)
    models = [W.n144_hh_cell(a) for a in amps]
    got32, _ = run_gpu(models, steps, precision=32, probes="soma")
    for i, mi in enumerate(models):
        r, _ = run_oracle(mi, steps, probe_cvs=[soma_cv(mi)])
        e = rmse(got32[:, i], r[:, 0])
        sg, sr = spike_times(got32[:, i], 0.025), spike_times(r[:, 0], 0.025)
        print(f"amp {amps[i]:.3f} nA: fp32 soma RMSE {e:.2e} mV, spikes {sg} / {sr}")
        assert e <= 0.05
        assert len(sg) == len(sr) and np.all(np.abs(sg - sr) <= 0.025 + 1e-9)


def test_dhs_voltage_integrator_matches_oracle(cuda):
    """dhs_voltage_mi355x: the voltage solve alone (gates and Ci untouched)
    through the registry, against the oracle's voltage-only step."""
    import torch
    from braincell_amd.cell import Cell
    from braincell_amd.quad import get_integrator
    from oracle.staggered import StaggeredOracle

    m = W.uneven_tree_cell(5, cmin=200, cmax=400)
    cell = Cell(m.morph, solver="dhs_voltage_mi355x", precision=64)
    cell.lower = lambda: m  # the lowered model as is
    cell.init_state(probes="all")
    step = get_integrator("dhs_voltage_mi355x")
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    dt = 0.025
    S0 = None
    for n in range(450):  # the clamp starts at 10 ms = step 400
        step(cell, n * dt, dt)
        o.voltage_step(n * dt, dt)
        if S0 is None:
            S0 = cell.runtime.S.clone()
    torch.cuda.synchronize()
    V, _ = cell.runtime.state_of_cell(0)
    err = np.abs(V - o.V).max()
    print(f"dhs_voltage: fp64 max |dV| after 450 steps {err:.2e} mV, V range {o.V.min():.2f}..{o.V.max():.2f}")
    assert err < 1e-9
    assert torch.equal(S0, cell.runtime.S)  # gates and Ci unchanged
