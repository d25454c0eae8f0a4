"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
the FP64 oracle) checked on the CPU: the oracle still reproduces them, the
product's host layout matches the reference-order rows, and each fixture is
pinned by what the mathematics fixes independently of the oracle (a dense
solve and charge balance for the toy tree, removable singularities and
closed-form limits for the rate tables)."""

import os

import numpy as np
import pytest

from tests.golden import make_golden as MG

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name))


# ---------------------------------------------------------------- c0 layout
def test_c0_layout_checksum_and_rows():
    g = _load("c0_layout.npz")
    assert g["parent_rows"].shape == (321,) and g["dynamic_rows"].shape == (256,)
    assert str(g["checksum"]) == MG.layout_checksum(g["parent_rows"], g["dynamic_rows"], g["peel"],
                                                    g["row_to_point_id"])
    # the oracle still produces the same tables
    now = MG.c0_layout()
    for k in ("parent_rows", "dynamic_rows", "row_to_point_id", "peel"):
        assert np.array_equal(now[k], g[k]), k
    for k in ("row_cap_uF", "diag_axial", "lowers", "uppers"):
        np.testing.assert_allclose(now[k], g[k], rtol=1e-14, atol=0)


def test_c0_layout_product_rows_match_fixture():
    from braincell_amd import workloads as W
    from braincell_amd.layout import discretize_cell

    g = _load("c0_layout.npz")
    m = W.single_tree_cell(0)
    geo = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    assert np.array_equal(geo.parent_rows, g["parent_rows"])
    assert np.array_equal(geo.dynamic_rows, g["dynamic_rows"])
    assert np.array_equal(geo.row_to_point_id, g["row_to_point_id"])
    assert np.array_equal(geo.peel[geo.row_to_point_id], g["peel"][g["row_to_point_id"]])
    for a, k in ((geo.row_cap_uF, "row_cap_uF"), (geo.diag_axial, "diag_axial"), (geo.lowers, "lowers"),
                 (geo.uppers, "uppers")):
        np.testing.assert_allclose(a, g[k], rtol=1e-13, atol=0)


def test_c0_layout_invariants():
    """Parents precede children, each row's diagonal is the negated sum of
    its off-diagonals (zero row sums of the axial operator), and the
    conductance is symmetric: cap[row] * lower[row] = cap[parent] * upper[row]."""
    g = _load("c0_layout.npz")
    par = g["parent_rows"]
    assert par[0] == -1 and np.all(par[1:] < np.arange(1, par.size))
    off = g["lowers"].copy()
    np.add.at(off, par[1:], g["uppers"][1:])
    np.testing.assert_allclose(g["diag_axial"], -off, rtol=1e-12)
    cap = g["row_cap_uF"]
    np.testing.assert_allclose(cap[1:] * g["lowers"][1:], cap[par[1:]] * g["uppers"][1:], rtol=1e-12)


# --------------------------------------------------------------- toy Hines
def _dense(parent, diag, lowers, uppers):
    n = parent.size
    A = np.diag(diag[:n].astype(np.float64))
    for r in range(1, n):
        A[r, parent[r]] = lowers[r]
        A[parent[r], r] = uppers[r]
    return A


def test_hines7_dense_solve_and_textbook_hines():
    from oracle.dhs import hines_sequential

    g = _load("hines7.npz")
    par = g["parent_rows"]
    assert par.size == 7
    n = 7
    A = _dense(par, g["diag"], g["lowers"], g["uppers"])
    x_dense = np.linalg.solve(A, g["rhs"][:n])
    np.testing.assert_allclose(g["x"][:n], x_dense, rtol=0, atol=1e-11)
    x_h = hines_sequential(par, g["diag"][:n].copy(), g["rhs"][:n].copy(), g["lowers"][:n], g["uppers"][:n])
    np.testing.assert_allclose(x_h, x_dense, rtol=0, atol=1e-11)
    np.testing.assert_array_equal(g["V1"], g["x"][g["dynamic_rows"]])


def test_hines7_charge_balance():
    """Implicit Euler on the passive cable conserves charge: the capacitive
    current of all CVs equals the leak current at V1 plus both clamps
    (axial currents cancel in the sum; boundary nodes hold no charge)."""
    from braincell_amd.layout import discretize_cell

    g = _load("hines7.npz")
    m = MG.toy7_model()
    geo = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    area_cm2 = geo.cv_area_um2 * 1e-8
    cap_uF = area_cm2 * 1.0
    dt = float(g["dt"])
    V0, V1 = g["V0"], g["V1"]
    i_cap_nA = cap_uF * (V1 - V0) / dt * 1e3  # uF * mV/ms = uA
    i_leak_nA = 0.0003 * area_cm2 * (-70.0 - V1) * 1e6  # S * mV = mA
    i_clamp_nA = 0.05 + 0.1
    assert abs(i_cap_nA.sum() - (i_leak_nA.sum() + i_clamp_nA)) < 1e-9


def test_hines7_oracle_reproduces():
    now = MG.hines7()
    g = _load("hines7.npz")
    for k in ("diag", "rhs", "lowers", "uppers", "x", "V1"):
        np.testing.assert_allclose(now[k], g[k], rtol=0, atol=1e-12)


# --------------------------------------------------------------- c0 traces
@pytest.mark.parametrize("name,maker", [("c0_trace.npz", lambda: MG.W.single_tree_cell(0)),
                                        ("ca_cell_trace.npz", MG.ca_cell)])
def test_trace_fixture_oracle_reproduces(name, maker):
    g = _load(name)
    now = MG.trace_fixture(maker())
    assert int(now["soma_cv"]) == int(g["soma_cv"])
    np.testing.assert_allclose(now["soma"], g["soma"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(now["all_cvs"], g["all_cvs"], rtol=0, atol=1e-10)
    # the clamp (10 ms = step 400) drives the soma up within the window
    s = g["soma"]
    assert abs(s[399] - s[0]) < 1e-3 and s[-1] - s[399] > 1.0


# -------------------------------------------------------------- rate tables
def test_rate_tables_oracle_reproduces():
    g = _load("rate_tables.npz")
    now = MG.rate_tables()
    for k in now:
        np.testing.assert_allclose(now[k], g[k], rtol=1e-14, atol=0, err_msg=k)


def test_rate_tables_pinned_by_closed_forms():
    g = _load("rate_tables.npz")
    V = g["V"]
    assert V[0] == -100.0 and V[-1] == 50.0 and V.size == 301

    def at(v):
        return int(np.argmin(np.abs(V - v)))

    # removable singularities of a t / (1 - exp(-t / k)) forms (V_sh = -50 mV)
    assert g["na_ap"][at(-37.0)] == pytest.approx(0.32 * 4.0, rel=1e-14)
    assert g["na_bp"][at(-10.0)] == pytest.approx(0.28 * 5.0, rel=1e-14)
    assert g["kdr_ap"][at(-35.0)] == pytest.approx(0.032 * 5.0, rel=1e-14)
    # the same forms, evaluated away from 0 by the textbook expression
    v = V[at(-60.0)] + 50.0
    assert g["na_ap"][at(-60.0)] == pytest.approx(0.32 * (13.0 - v) / (np.exp((13.0 - v) / 4.0) - 1.0), rel=1e-13)
    assert g["kdr_bp"][at(0.0)] == pytest.approx(0.5 * np.exp(-(50.0 - 10.0) / 40.0), rel=1e-14)
    # half-activation points of the logistic steady states
    assert g["cal_pinf"][at(-10.0)] == pytest.approx(0.5, rel=1e-14)
    assert g["cal_qinf"][at(-25.0)] == pytest.approx(0.5, rel=1e-14)
    assert g["hcn_pinf"][at(-75.0)] == pytest.approx(0.5, rel=1e-14)
    assert g["cat_pinf"][at(-62.0)] == pytest.approx(0.5, rel=1e-14)  # V_sh = -3
    # CaL tau_q peaks at V = -40 mV with 300 + 100 / 2
    assert g["cal_qtau"][at(-40.0)] == pytest.approx(350.0, rel=1e-14)
    assert np.argmax(g["cal_qtau"]) == at(-40.0)
    # steady states in [0, 1] (steep ones saturate in FP64) and monotone
    for k in ("cal_pinf", "cat_pinf", "hcn_pinf", "cal_qinf", "cat_qinf"):
        assert np.all((g[k] > 0) & (g[k] <= 1)), k
    assert np.all(np.diff(g["cal_pinf"]) >= 0) and np.all(np.diff(g["hcn_pinf"]) <= 0)
    # AHP: alpha = 48 Ci^2, beta = 0.09; the Nernst E_Ca falls by
    # RT/(2F) ln 10 per decade of Ci
    Ci = g["Ci"]
    np.testing.assert_allclose(g["ahp_ap"], 48.0 * Ci ** 2, rtol=1e-14)
    assert np.all(g["ahp_bp"] == 0.09)
    dec = 8.314462618 * (36.0 + 273.15) / (2 * 96485.33212) * np.log(10.0) * 1e3
    np.testing.assert_allclose(np.diff(g["e_ca"][::20]), -dec, rtol=1e-12)


def test_product_steady_states_match_rate_tables():
    """The host's reset_state initialisers (braincell_amd/rates.py) equal
    alpha / (alpha + beta) and x_inf read from the tables."""
    from braincell_amd import rates

    g = _load("rate_tables.npz")
    V = g["V"]
    m, h = rates.na_steady(V, -50.0)
    np.testing.assert_allclose(m, g["na_ap"] / (g["na_ap"] + g["na_bp"]), rtol=1e-13)
    np.testing.assert_allclose(h, g["na_aq"] / (g["na_aq"] + g["na_bq"]), rtol=1e-13)
    np.testing.assert_allclose(rates.kdr_steady(V, -50.0), g["kdr_ap"] / (g["kdr_ap"] + g["kdr_bp"]), rtol=1e-13)
    p, q = rates.cal_steady(V, 0.0)
    np.testing.assert_allclose(p, g["cal_pinf"], rtol=1e-14)
    np.testing.assert_allclose(q, g["cal_qinf"], rtol=1e-14)
    p, q = rates.cat_steady(V, -3.0)
    np.testing.assert_allclose(p, g["cat_pinf"], rtol=1e-14)
    np.testing.assert_allclose(q, g["cat_qinf"], rtol=1e-14)
    np.testing.assert_allclose(rates.hcn_steady(V), g["hcn_pinf"], rtol=1e-14)
    np.testing.assert_allclose(rates.ahp_steady(g["Ci"], 48.0, 0.09, 2.0),
                               g["ahp_ap"] / (g["ahp_ap"] + g["ahp_bp"]), rtol=1e-14)
