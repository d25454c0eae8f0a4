"""Edge cases of the HIP paths against the oracle: launches of zero steps,
ragged extremes in one workgroup (a one-compartment soma beside a
1700-compartment tree), the largest cell one workgroup holds (4000
compartments, 16 per thread, ~147 KB of LDS in FP32), and point
populations that do not fill a workgroup."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import rmse, run_gpu, run_oracle, soma_cv, spike_times

pytestmark = pytest.mark.gpu


def _soma_only_cell(amp=0.2):
    """One branch, one CV: the soma of the benchmark trees with Na/KDR/leak
    and a clamp from 1 ms."""
    from braincell_amd.model import CellModel
    from braincell_amd.morph import Morphology

    m = Morphology.from_root(W._soma())
    g = W.BASE_DENSITIES
    return CellModel(
        morph=m, cv_per_branch=1, cm=1.0, ra=100.0, V0=-70.0,
        channels={"Na_Ba2002": {"g_branch": np.full(1, g["Na_Ba2002"]), "params": {}},
                  "KDR_Ba2002": {"g_branch": np.full(1, g["KDR_Ba2002"]), "params": {}},
                  "IL": {"g_branch": np.full(1, g["IL"]), "params": {"E": -70.0}}},
        clamps=[(0, 0.5, 1.0, W.CLAMP_FOREVER, amp)],
    ).validate()


def test_zero_step_launches_change_nothing(cuda):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime
    from braincell_amd.single import SingleCompartment

    L = StaggerLayout([W.uneven_tree_cell(i, cmin=200, cmax=400) for i in range(2)], probes="soma")
    rt = StaggerRuntime(L, 0.025, precision=32)
    rt.run(40)
    torch.cuda.synchronize()
    V, S, Vb = rt.V.clone(), rt.S.clone(), rt.Vbnd.clone()
    tr = rt.run(0)
    torch.cuda.synchronize()
    rt.check_status()
    assert tr.shape == (0, L.n_probes)
    assert torch.equal(rt.V, V) and torch.equal(rt.S, S) and torch.equal(rt.Vbnd, Vb)

    pop = W.htc_population(300, seed=1)
    channels = {nm: dict(W.HTC_CHANNELS[nm], g_max=pop["g"][nm]) for nm in pop["names"]}
    sc = SingleCompartment(300, channels=channels, E_Na=W.HTC_IONS["E_Na"], E_K=W.HTC_IONS["E_K"],
                           ca=W.HTC_IONS["ca"], V0=pop["V0"], precision=32)
    sc.init_state()
    sc.set_input(pop["I_dc_nA"])
    sc.advance(0.0, 0.025, 40)
    torch.cuda.synchronize()
    V, S = sc.V.clone(), sc._dev["S"].clone()
    sc.advance(1.0, 0.025, 0)
    torch.cuda.synchronize()
    assert torch.equal(sc.V, V) and torch.equal(sc._dev["S"], S)


def test_ragged_extremes_in_one_pack_fp64(cuda):
    """A 1-CV soma, the 256-CV benchmark tree and a 1700-CV tree share one
    workgroup (cells of 3, 321 and ~2100 DHS rows); every CV of every cell
    follows the oracle."""
    models = [_soma_only_cell(), W.single_tree_cell(0), W.uneven_tree_cell(5, cmin=1700, cmax=1700)]
    steps = 800  # the trees' clamps start at 10 ms (step 400)
    got, rt = run_gpu(models, steps, precision=64, probes="all", threads=256, cv_per_thread=8, rows_per_pack=2500)
    L = rt.L
    assert L.n_packs == 1, "the three cells must share one workgroup"
    assert L.lds_bytes(64) <= 160 * 1024
    for i, m in enumerate(models):
        ref, _ = run_oracle(m, steps)
        c0 = int(L.cell_probe0[i])
        blk = got[:, c0:c0 + ref.shape[1]]
        err = np.abs(blk - ref).max()
        print(f"cell {i} ({ref.shape[1]} CVs): fp64 max |dV| over all CVs {err:.2e} mV")
        assert err < 1e-5
    # the lone soma fires under its clamp: the path is not trivially at rest
    assert len(spike_times(got[:, int(L.cell_probe0[0])], 0.025)) > 0


def test_largest_cell_one_workgroup_holds_fp32(cuda):
    """4000 CVs (~5000 DHS rows) at 16 CVs per thread: FP32 soma within
    0.05 mV of the FP64 oracle, spikes at the same steps."""
    m = W.uneven_tree_cell(11, cmin=4000, cmax=4000)
    steps, dt = 1200, 0.025
    got, rt = run_gpu([m], steps, precision=32, probes="soma", threads=256, cv_per_thread=16)
    assert rt.L.max_rows > 4500 and rt.L.lds_bytes(32) <= 160 * 1024
    ref, _ = run_oracle(m, steps, probe_cvs=[soma_cv(m)])
    e = rmse(got[:, 0], ref[:, 0])
    st_g, st_r = spike_times(got[:, 0], dt), spike_times(ref[:, 0], dt)
    print(f"4000-CV cell: {rt.L.max_rows} rows, fp32 soma RMSE {e:.2e} mV, spikes {len(st_g)} / {len(st_r)}")
    assert e <= 0.05
    assert len(st_g) == len(st_r)
    if len(st_r):
        assert np.max(np.abs(st_g - st_r)) <= dt + 1e-9


@pytest.mark.parametrize("n", [1, 257])
def test_point_populations_that_do_not_fill_a_workgroup_fp64(cuda, n):
    from tests.test_gpu_point import _oracle, _run

    steps = 2000
    got, pop, ch, _ = _run(n, steps, 64, n)
    ref = _oracle(pop, ch, n, steps)
    err = np.abs(got - ref).max()
    print(f"{n} neurons: fp64 max |dV| {err:.2e} mV")
    assert err < 1e-6
