"""Pins the FP64 oracle against the reference's known answers, closed forms,
dense linear algebra, library routines and finite differences (CPU only)."""

import numpy as np
import pytest
import scipy.constants as sc
import scipy.linalg as sla
from scipy.special import exprel

from braincell_amd import workloads as W
from braincell_amd.morph import Branch, Morphology
from braincell_amd.model import CellModel
from oracle import channels as ch
from oracle import discretize as OD
from oracle import dhs as OH
from oracle.staggered import StaggeredOracle


# ------------------------------------------------ reference known answers
def test_peel_levels_known_answer():
    # _compute/scheduling_test.py:30-41
    lv = OD.compute_peel_levels(np.asarray([3, 3, 0, -1]), ((2,), (), (), (0, 1)))
    assert lv.tolist() == [1, 0, 0, 2]


def test_peel_levels_isolated_and_cycle():
    assert OD.compute_peel_levels(np.asarray([-1, -1]), ((), ())).tolist() == [0, 0]
    with pytest.raises(ValueError, match="cycle"):
        OD.compute_peel_levels(np.asarray([1, 0]), ((1,), (0,)))


def test_backsub_indices_known_answers():
    # quad/_staggered_test.py:182-199
    idx = OD.build_backsub_indices(np.asarray([1, 1]), n_nodes=1)
    assert idx.shape[1] == 2 and int(idx[0, 0]) == 1
    idx = OD.build_backsub_indices(np.asarray([1, 2, 3, 3]), n_nodes=3)
    assert idx[0].tolist() == [1, 2, 3, 3]
    assert int(idx[1, 0]) == 2


@pytest.mark.parametrize("maker", [lambda: W.single_tree_cell(0), lambda: W.uneven_tree_cell(5, cmin=200, cmax=400)])
def test_scheduling_invariants(maker):
    # _compute/scheduling_test.py:59-104
    m = maker()
    _, _, tree, sched, st = OD.discretize(m.morph, 4)
    rows = sorted(int(r) for g in sched.groups for r in g.tolist())
    assert rows == list(range(tree.n_nodes))
    pairs = {tuple(e) for e in st.edges.tolist()}
    assert len(pairs) == st.edges.shape[0] == tree.n_nodes - 1
    # every level chunk is free of parent/child pairs inside it
    for g in sched.groups:
        gs = set(g.tolist())
        assert not any(int(sched.parent_rows[r]) in gs for r in g.tolist())


def test_node_count_formula():
    m = W.single_tree_cell(0)
    _, _, tree, _, st = OD.discretize(m.morph, 4)
    assert tree.n_nodes == 256 + 64 + 1 == st.n_point


# ------------------------------------------------------- geometry closed forms
def test_cylinder_and_frustum_closed_forms():
    L, r = 40.0, 2.0
    br = Branch.from_lengths(lengths=[L], radii=[r, r], type="dendrite")
    cvs, _ = OD.build_cvs(Morphology.from_root(br), 4, 1.0, 100.0)
    for cv in cvs:
        assert cv.area_um2 == pytest.approx(2 * np.pi * r * L / 4, rel=1e-14)
        half = 100.0 * (L / 8 * 1e-4) / (np.pi * (r * 1e-4) ** 2)
        assert cv.r_axial_prox == pytest.approx(half, rel=1e-13)
        assert cv.r_axial_dist == pytest.approx(half, rel=1e-13)
    r0, r1 = 2.0, 1.0
    br = Branch.from_lengths(lengths=[L], radii=[r0, r1], type="dendrite")
    cvs, _ = OD.build_cvs(Morphology.from_root(br), 1, 1.0, 100.0)
    slant = np.hypot(L, r1 - r0)
    assert cvs[0].area_um2 == pytest.approx(np.pi * (r0 + r1) * slant, rel=1e-14)
    # resistance of a linear taper = Ra L / (pi r0 r1) (exact integral of dx / (pi r(x)^2))
    rm = 0.5 * (r0 + r1)
    tot = 100.0 * (L * 1e-4) / (np.pi * r0 * r1 * 1e-8)
    assert cvs[0].r_axial_prox + cvs[0].r_axial_dist == pytest.approx(tot, rel=1e-13)
    assert cvs[0].r_axial_prox == pytest.approx(100.0 * (L / 2 * 1e-4) / (np.pi * r0 * rm * 1e-8), rel=1e-13)


# ------------------------------------------------------------ DHS linear algebra
def _random_system(st, rng, dt=0.025):
    C = st.dynamic_rows.shape[0]
    V = rng.uniform(-80, 0, (1, C))
    lin = -rng.uniform(0, 2, (1, C))
    const = rng.uniform(-10, 10, (1, C))
    return V, lin, const, dt


@pytest.mark.parametrize("seed", [0, 1])
def test_dhs_solve_equals_dense_solve(seed):
    m = W.single_tree_cell(0) if seed == 0 else W.uneven_tree_cell(3, cmin=200, cmax=300)
    _, _, _, _, st = OD.discretize(m.morph, 4)
    rng = np.random.default_rng(seed)
    V, lin, const, dt = _random_system(st, rng)
    got = OH.dhs_solve(st, V, lin, const, dt)[0]
    d, s, lo, up = OH.numeric_state(st, V, lin, const, dt)
    M = OH.assemble_dense(st, d[0], lo, up)
    x = np.linalg.solve(M, s[0, :st.n_point])
    np.testing.assert_allclose(got, x[st.dynamic_rows], rtol=0, atol=1e-9)
    # the textbook sequential Hines sweep gives the same solution
    xh = OH.hines_sequential(st.parent_rows, d[0, :st.n_point], s[0, :st.n_point], lo[:st.n_point], up[:st.n_point])
    np.testing.assert_allclose(xh, x, rtol=0, atol=1e-9)


def test_dhs_equals_schur_reduced_cv_operator():
    # _staggered.py:393-413: the boundary rows are exact constraints
    m = W.uneven_tree_cell(4, cmin=200, cmax=260)
    _, _, _, _, st = OD.discretize(m.morph, 4)
    rng = np.random.default_rng(3)
    V, lin, const, dt = _random_system(st, rng)
    got = OH.dhs_solve(st, V, lin, const, dt)[0]
    Ared = OD.schur_cv_operator(st)
    M = np.eye(Ared.shape[0]) + dt * Ared - dt * np.diag(lin[0])
    x = np.linalg.solve(M, V[0] + dt * const[0])
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-9)


def test_unbranched_cable_matches_solve_banded():
    br = Branch.from_lengths(lengths=[300.0], radii=[1.0, 0.6], type="dendrite")
    _, _, _, _, st = OD.discretize(Morphology.from_root(br), 12)
    rng = np.random.default_rng(7)
    V, lin, const, dt = _random_system(st, rng)
    got = OH.dhs_solve(st, V, lin, const, dt)[0]
    Ared = OD.schur_cv_operator(st)
    M = np.eye(12) + dt * Ared - dt * np.diag(lin[0])
    ab = np.zeros((3, 12))
    ab[0, 1:] = np.diag(M, 1)
    ab[1] = np.diag(M)
    ab[2, :-1] = np.diag(M, -1)
    assert np.allclose(M, np.triu(np.tril(M, 1), -1))  # an unbranched cable is tridiagonal
    x = sla.solve_banded((1, 1), ab, V[0] + dt * const[0])
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-10)


def test_passive_step_conserves_charge_and_reaches_rest():
    # with leak only, V relaxes to E_leak everywhere; the axial operator has zero row sums
    m = W.single_tree_cell(0)
    m.channels = {"IL": m.channels["IL"]}
    m.clamps = []
    m.V0 = -50.0
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    C_row = o.st.row_cap_uF
    assert np.allclose((C_row[:, None] * o.st.axial).sum(axis=0), 0, atol=1e-12 * np.abs(C_row[:, None] * o.st.axial).max())
    tr = o.run(0.025, 4000, probe_cvs=[0, 100, 255])
    assert np.allclose(tr[-1], -70.0, atol=1e-3)


# ----------------------------------------------------------- channel math
def test_exp_euler_linear_decay_known_answer():
    # quad/_exp_euler_test.py:113-120: dx/dt = -x / tau, one step = exp(-dt/tau)
    y = ch.exp_euler(1.0, -1.0 / 1.0, -1.0 / 1.0, 0.01)
    assert y == pytest.approx(np.exp(-0.01), abs=1e-12)


def test_gate_update_closed_form_and_fixed_point():
    V = np.linspace(-100, 50, 301)
    ap, bp, aq, bq = ch.na_ba2002(V)
    x0 = np.full_like(V, 0.3)
    dt = 0.025
    got = ch.ab_update(x0, ap, bp, 1.0, dt)
    xinf = ap / (ap + bp)
    np.testing.assert_allclose(got, xinf + (x0 - xinf) * np.exp(-dt * (ap + bp)), rtol=0, atol=1e-12)
    np.testing.assert_allclose(ch.ab_update(xinf, ap, bp, 1.0, dt), xinf, rtol=0, atol=1e-14)
    pi, pt, qi, qt = ch.cal_is2008(V)
    got = ch.it_update(x0, pi, pt, 2.0, dt)
    np.testing.assert_allclose(got, pi + (x0 - pi) * np.exp(-dt * 2.0 / pt), rtol=0, atol=1e-12)


def test_rate_singularities_are_the_limits():
    # exprel removes the 0/0 at the Boltzmann midpoints: V' = 13 gives 0.32 * 4
    ap, bp, _, _ = ch.na_ba2002(np.asarray([-50.0 + 13.0]))
    assert ap[0] == pytest.approx(1.28, abs=1e-14)
    ap, bp, _, _ = ch.na_ba2002(np.asarray([-50.0 + 40.0]))
    assert bp[0] == pytest.approx(1.4, abs=1e-14)
    a, _ = ch.kdr_ba2002(np.asarray([-50.0 + 15.0]))
    assert a[0] == pytest.approx(0.16, abs=1e-14)
    assert exprel(0.0) == 1.0


def test_cat_tau_branch_is_continuous_enough():
    # calcium.py:369-373: both branches of tau_q evaluated across V' = -80
    _, _, _, qt = ch.cat_hm1992(np.asarray([-77.0 - 1e-9, -77.0]))
    assert abs(qt[0] - qt[1]) / qt[1] < 0.02


def test_constants_are_codata2018_and_nernst():
    # CODATA 2018 publishes F = 96485.33212...; N_A e (exact SI) = 96485.33212331
    assert ch.FARADAY == pytest.approx(sc.physical_constants["Faraday constant"][0], rel=1e-10)
    assert ch.GAS_CONSTANT == pytest.approx(sc.R, rel=1e-10)
    assert ch.nernst(2.0, 2.0) == 0.0
    # z F E = R T ln(Co/Ci), independently with scipy constants
    E = ch.nernst(2.4e-4, 2.0, 309.15, 2.0)
    assert E == pytest.approx(1e3 * sc.R * 309.15 / (2 * sc.physical_constants["Faraday constant"][0])
                              * np.log(2.0 / 2.4e-4), rel=1e-10)


def test_ca_drive_units():
    # 1 uA/cm^2 = 1e-2 A/m^2 into a 1 um shell: I / (2 F d) mol/(m^3 s) = mM/s
    F = sc.physical_constants["Faraday constant"][0]
    expect = 1e-2 / (2 * F * 1e-6) / 1e3
    assert ch.ca_detailed_drive(1.0, 1.0) == pytest.approx(expect, rel=1e-12)
    assert ch.ca_detailed_drive(-1.0, 1.0) == 0.0


def test_linearisation_matches_finite_difference():
    m = W.uneven_tree_cell(2, cmin=200, cmax=300)
    o = StaggeredOracle(m)
    o.reset_state(-55.0)
    rng = np.random.default_rng(1)
    for name in o.gates:
        o.gates[name] = [rng.uniform(0.05, 0.95, o.C) for _ in o.gates[name]]
    o.Ci = rng.uniform(1e-4, 1e-3, o.C)
    V = rng.uniform(-80, 20, o.C)
    I, G = o.membrane_current(V, o.Ci)
    h = 1e-5
    Ip, _ = o.membrane_current(V + h, o.Ci)
    Im, _ = o.membrane_current(V - h, o.Ci)
    np.testing.assert_allclose((Ip - Im) / (2 * h), -G, rtol=1e-7, atol=1e-9)


def test_gate_lambda_matches_finite_difference():
    # ind_exp_euler uses lam = d f / d x with the other states frozen
    V = np.linspace(-90, 40, 53)
    ap, bp, _, _ = ch.na_ba2002(V)
    f = lambda x: 1.0 * (ap * (1 - x) - bp * x)
    h = 1e-6
    np.testing.assert_allclose((f(0.4 + h) - f(0.4 - h)) / (2 * h), -(ap + bp), rtol=1e-7)


def test_family_order_ca_before_gates():
    # AHP-free check of the staggered order: Ci uses the step-start current and
    # gates see V_{n+1}; swapping the order changes the result measurably
    m = W.uneven_tree_cell(1, cmin=200, cmax=300)
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    for n in range(500):
        o.step(n * 0.025, 0.025)
    Vn, Ci_n = o.V.copy(), o.Ci.copy()
    I_ca = o.ca_current(o.V, o.Ci)
    o.step(500 * 0.025, 0.025)
    fc = ch.ca_detailed_drive(I_ca, 1.0) + (2.4e-4 - Ci_n) / 5.0
    np.testing.assert_allclose(o.Ci, ch.exp_euler(Ci_n, fc, -1 / 5.0 * np.ones_like(Ci_n), 0.025), rtol=1e-13)
