"""The paint/place ``Cell`` front end and the registry-selected solver.

A braincell user builds cells with ``paint``/``place`` and picks the solver by
name (``Cell(solver=...)``, ``_multi_compartment/cell.py:242, 2771-2776``);
these tests check that the front end lowers to exactly the model the
benchmarks use and that the registered step advances it like a full run.
"""

import numpy as np
import pytest

from braincell_amd import environ
from braincell_amd import workloads as W
from braincell_amd.cell import Cell, CellBatch, CVPerBranch
from braincell_amd.filter import AllRegion, RootLocation, branch_in
from braincell_amd.mech import CableProperty, Channel, CurrentClamp, Ion
from braincell_amd.quad import get_integrator, staggered_mi355x_step
from braincell_amd.runtime import StaggerLayout


def c0_cell(**kw) -> Cell:
    """Config 0 written the way a braincell user writes it."""
    cell = Cell(W.binary_tree_morphology(0), cv_policy=CVPerBranch(4), **kw)
    cell.paint(AllRegion(), CableProperty(resting_potential=-70.0, membrane_capacitance=1.0, axial_resistivity=100.0),
               Channel("IL", g_max=3e-4, E=-70.0))
    cell.paint(branch_in("type", "soma"), Channel("INa_Ba2002", g_max=0.12), Channel("KDR_Ba2002", g_max=0.036))
    cell.place(RootLocation(0.5), CurrentClamp(delay=10.0, durations=[50.0], amplitudes=[0.2]))
    return cell


def test_paint_place_lowers_to_the_c0_model():
    got = c0_cell().lower()
    ref = W.single_tree_cell(0)
    assert got.V0 == ref.V0
    assert np.array_equal(got.per_branch(got.cm), ref.per_branch(ref.cm))
    assert np.array_equal(got.per_branch(got.ra), ref.per_branch(ref.ra))
    assert set(got.channels) == set(ref.channels)
    for name in ref.channels:
        assert np.array_equal(got.channels[name]["g_branch"], ref.channels[name]["g_branch"]), name
    assert got.clamps == ref.clamps
    assert (got.E_Na, got.E_K) == (ref.E_Na, ref.E_K)
    # the flat layout (every table the kernel reads) is identical too
    a, b = StaggerLayout([got]), StaggerLayout([ref])
    for k in ("cv_info", "bnd_info", "chain_off", "chain_top", "kid_off", "kids", "level_off", "slot_lo",
              "slot_up"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for x, y in zip(a.dt_tables(0.025), b.dt_tables(0.025)):
        assert np.array_equal(np.asarray(x), np.asarray(y))


def test_later_paint_rules_win_and_ions_are_painted():
    cell = Cell(W.binary_tree_morphology(0), cv_policy=CVPerBranch(2))
    cell.paint(AllRegion(), Channel("IL", g_max=1e-4))
    cell.paint(branch_in("index", [1, 2]), Channel("IL", g_max=5e-4))
    cell.paint(branch_in("type", "dendrite"), Channel("CaL_IS2008", g_max=0.002), Ion("CalciumDetailed", tau=7.0))
    m = cell.lower()
    g = m.channels["IL"]["g_branch"]
    assert g[0] == 1e-4 and g[1] == 5e-4 and g[2] == 5e-4 and np.all(g[3:] == 1e-4)
    assert m.ca is not None and m.ca["tau"] == 7.0
    assert not m.ca["mask_branch"][0] and np.all(m.ca["mask_branch"][1:])


def test_paint_and_place_reject_wrong_mechanisms():
    cell = Cell(W.binary_tree_morphology(0))
    with pytest.raises(TypeError):
        cell.paint(AllRegion(), CurrentClamp(delay=1.0))
    with pytest.raises(TypeError):
        cell.place(RootLocation(0.5), Channel("IL"))
    with pytest.raises(ValueError):
        Channel("NoSuchChannel")
    with pytest.raises(TypeError):
        Channel("IL", g_maxx=1.0)


def test_solver_is_resolved_by_name_or_alias_at_construction():
    assert c0_cell().solver is staggered_mi355x_step
    assert c0_cell(solver="stagger_mi355x").solver is staggered_mi355x_step
    assert get_integrator(staggered_mi355x_step) is staggered_mi355x_step
    with pytest.raises(ValueError, match="Unknown integrator"):
        c0_cell(solver="staggered_mi355")
    with pytest.raises(TypeError):
        c0_cell(solver=3)


def test_solver_cannot_be_reassigned_after_init_state():
    cell = c0_cell()
    cell.solver = "stagger_mi355x"  # allowed before init
    cell.init_state()
    with pytest.raises(RuntimeError):
        cell.solver = "staggered_mi355x"


def test_step_rejects_wrong_targets_and_update_orders():
    with pytest.raises(TypeError):
        staggered_mi355x_step(object())
    cell = c0_cell(ion_channel_update_order="sideways")
    with environ.context(dt=0.025, t=0.0), pytest.raises(ValueError):
        cell.update()


def test_cell_batch_accepts_cells_and_models():
    b = CellBatch([c0_cell(), W.single_tree_cell(0)], device="cpu")
    assert len(b.layout.cells) == 2 and b.layout.n_cv_total == 2 * 256


@pytest.mark.gpu
def test_cell_run_and_registry_steps_match_the_oracle(cuda):
    from tests.helpers import run_oracle

    n = 400
    ref, _ = run_oracle(W.single_tree_cell(0), n)  # [n, C] FP64, every CV
    cell = c0_cell(precision=64)
    res = cell.run(dt=0.025, duration=n * 0.025)
    soma = cell._layout.cells[0].probes[0]
    assert res.traces["V"].shape == (n, 1)
    assert np.max(np.abs(res.traces["V"][:, 0] - ref[:, soma])) < 1e-8
    # the same trajectory one registered step at a time (t, dt from environ)
    stepper = c0_cell(precision=64)
    vs = []
    with environ.context(dt=0.025):
        for i in range(60):
            with environ.context(t=i * 0.025):
                stepper.update()
            vs.append(float(stepper.last_trace[0, 0]))
    assert np.max(np.abs(np.asarray(vs) - res.traces["V"][:60, 0])) < 1e-10
    # a run continues from where the steps stopped (clamp windows follow t)
    more = stepper.run(dt=0.025, duration=(n - 60) * 0.025)
    assert np.max(np.abs(more.traces["V"][:, 0] - res.traces["V"][60:, 0])) < 1e-10


def _probe_cell():
    from braincell_amd.filter import BranchLocation, RootLocation
    from braincell_amd.mech import CurrentProbe, MechanismProbe, StateProbe

    m = W.uneven_tree_cell(7, cmin=200, cmax=300)
    cell = Cell(m.morph, cv_policy=CVPerBranch(4))
    cell.lower = lambda: m  # the workload's lowered model as is
    cell.place(RootLocation(0.5), StateProbe("v_soma"), MechanismProbe("na_p", "Na_Ba2002", "p"))
    cell.place(BranchLocation(3, 0.4), StateProbe("v_dend"), MechanismProbe("ci", "CalciumDetailed", "C"),
               MechanismProbe("cal_q", "CaL_IS2008", "q"), CurrentProbe("i_cal", mechanism="CaL_IS2008"),
               CurrentProbe("i_ca", ion="Ca"), CurrentProbe("i_kdr", mechanism="KDR_Ba2002"))
    cell.place(BranchLocation(2, 1.0), StateProbe("v_tip"))
    return cell, m


def test_probe_placement_and_locations():
    from braincell_amd.mech import CurrentProbe, StateProbe

    cell, m = _probe_cell()
    pts = {p.name: cvs for p, cvs in cell._probe_points()}
    assert pts["v_soma"] == pts["na_p"] and len(pts["v_soma"]) == 1
    # branch 2, x = 1: a boundary node, sampled at the branch's last CV
    first = int(np.sum(m.n_cv_branch()[:2]))
    assert pts["v_tip"] == [first + m.n_cv_branch()[2] - 1]
    cell.init_state()  # host layout only: the probe CVs become trace columns
    assert sorted(int(c) for c in cell._layout.cells[0].probes) == sorted({c for v in pts.values() for c in v})
    with pytest.raises(ValueError, match="unique"):
        cell.place(RootLocation(0.5), StateProbe("v_soma"))
    with pytest.raises(ValueError):
        CurrentProbe("x")
    with pytest.raises(ValueError):
        StateProbe("x", field="m")


@pytest.mark.gpu
def test_probes_match_the_oracle(cuda):
    """Voltage, gate, pool and current probes of Cell.run against the FP64
    oracle's values at the same CVs after every step."""
    from oracle.staggered import StaggeredOracle

    cell, m = _probe_cell()
    cell.precision = 64
    cell.init_state()
    n = 480
    res = cell.run(dt=0.025, duration=n * 0.025)
    assert sorted(res.traces) == list(res.traces)
    pts = {p.name: cvs for p, cvs in cell._probe_points()}
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    ref = {k: np.zeros((n, len(v))) for k, v in pts.items()}
    for s in range(n):
        o.step(s * 0.025, 0.025)
        for k, cvs in pts.items():
            if k.startswith("v_"):
                val = o.V[cvs]
            elif k == "na_p":
                val = o.gates["Na_Ba2002"][0][cvs]
            elif k == "cal_q":
                val = o.gates["CaL_IS2008"][1][cvs]
            elif k == "ci":
                val = o.Ci[cvs]
            elif k == "i_kdr":
                val = (o._g_eff("KDR_Ba2002") * (o._E("KDR_Ba2002", o.Ci) - o.V))[cvs]
            else:  # i_cal, i_ca (CaL is the only Ca channel)
                val = (o._g_eff("CaL_IS2008") * (o._E("CaL_IS2008", o.Ci) - o.V))[cvs]
            ref[k][s] = val
    for k in pts:
        got = res.traces[k]
        scale = np.abs(ref[k]).max()
        err = np.abs(got - ref[k]).max() / scale
        print(f"{k}: max rel err {err:.2e} (scale {scale:.3g}, range {np.ptp(ref[k]):.3g})")
        assert got.shape == (n, len(pts[k]))
        assert np.ptp(ref[k]) > 0.0, k  # the probed quantity moves during the run
        assert err < 1e-8, k
