"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference cannot be run here (its JAX stack is absent), so the vectors
are produced by the oracle's restatement and pinned, in the tests that read
them, by what the mathematics fixes (dense solves, removable singularities,
steady states) as well as by the oracle itself.  Fixtures:

- ``c0_layout.npz``: the 256-CV single-tree cell's DHS rows in reference order
  (``quad/_staggered.py:323-390``, ``_compute/scheduling.py:45-213``): parent
  rows, CV -> row map, peel levels, row capacitance and axial coefficients,
  plus a SHA-256 checksum over the integer tables.
- ``c0_trace.npz``: FP64 soma trace of the same cell over its first 450 steps
  (dt 0.025 ms) and every CV over steps 400-449 (the 0.2 nA clamp starts at
  10 ms = step 400).
- ``ca_cell_trace.npz``: the same for a 248-CV Ca-bearing cell (seed 0,
  Na/KDR/leak everywhere, CaL and a CalciumDetailed pool on the dendrites).
- ``hines7.npz``: one implicit step of a 7-row toy tree (soma and two
  dendrites, one CV each; P = C + B + 1 = 7) with a midpoint clamp and a
  clamp on a terminal boundary node: the assembled rows, the row solution
  and the new CV voltages.
- ``rate_tables.npz``: per-channel rate functions on V = -100 ... +50 mV in
  0.5 mV steps, AHP rates and E_Ca on a Ci grid.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from braincell_amd import workloads as W  # noqa: E402
from braincell_amd.model import CellModel  # noqa: E402
from braincell_amd.morph import Branch, Morphology  # noqa: E402
from oracle import channels as ch  # noqa: E402
from oracle import discretize as OD  # noqa: E402
from oracle.dhs import dhs_solve_rows, numeric_state  # noqa: E402
from oracle.staggered import StaggeredOracle  # noqa: E402

DT = 0.025
V_GRID = np.arange(-100.0, 50.0 + 0.25, 0.5)
CI_GRID = np.geomspace(1e-5, 1e-1, 81)


def layout_checksum(parent_rows, dynamic_rows, peel, row_to_point_id) -> str:
    h = hashlib.sha256()
    for a in (parent_rows, dynamic_rows, peel, row_to_point_id):
        h.update(np.ascontiguousarray(a, dtype="<i8").tobytes())
    return h.hexdigest()


def c0_layout():
    m = W.single_tree_cell(0)
    _, _, tree, sched, st = OD.discretize(m.morph, 4, 1.0, 100.0)
    parent, children = OD.node_parent_children(tree)
    peel = OD.compute_peel_levels(parent, children)
    out = {
        "parent_rows": st.parent_rows.astype(np.int64),
        "dynamic_rows": st.dynamic_rows.astype(np.int64),
        "row_to_point_id": np.asarray(sched.row_to_point_id, dtype=np.int64),
        "peel": np.asarray(peel, dtype=np.int64),
        "row_cap_uF": st.row_cap_uF,
        "diag_axial": st.diag,
        "lowers": st.lowers,
        "uppers": st.uppers,
    }
    out["checksum"] = np.array(layout_checksum(out["parent_rows"], out["dynamic_rows"], out["peel"],
                                               out["row_to_point_id"]))
    return out


def trace_fixture(m, n_steps=450, window=50):
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    full = o.run(DT, n_steps)
    soma = _soma_cv(m)
    return {"soma_cv": np.int64(soma), "soma": full[:, soma].copy(), "window_start": np.int64(n_steps - window),
            "all_cvs": full[n_steps - window:].copy(), "dt": np.float64(DT)}


def _soma_cv(m):
    from braincell_amd.layout import discretize_cell

    g = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    return g.locate(0, 0.5, m.n_cv_branch())[1]


def toy7_model() -> CellModel:
    soma = Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma")
    morph = Morphology.from_root(soma)
    morph.attach(0, Branch.from_lengths(lengths=[100.0], radii=[1.0, 0.5], type="dendrite"))
    morph.attach(0, Branch.from_lengths(lengths=[50.0], radii=[0.75, 0.375], type="dendrite"))
    return CellModel(morph=morph, cv_per_branch=1, cm=1.0, ra=100.0, V0=-70.0,
                     channels={"IL": {"g_branch": np.full(3, 0.0003), "params": {"E": -70.0}}},
                     clamps=[(0, 0.5, 0.0, 1.0e9, 0.05), (1, 1.0, 0.0, 1.0e9, 0.1)]).validate()


TOY7_V0 = np.array([-70.0, -52.5, -81.25])


def hines7():
    m = toy7_model()
    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    o.V = TOY7_V0.copy()
    st = o.st
    V = o.V
    I, G = o.membrane_current(V, None)
    I = I + o.clamp_density(0.0)
    lin = -G / o.cm
    const = I / o.cm - V * lin
    d, s, lo, up = numeric_state(st, V[None, :], lin[None, :], const[None, :], DT, o.edge_current(0.0))
    x = dhs_solve_rows(st, d.copy(), s.copy(), lo, up)[0]
    o.step(0.0, DT)
    return {"V0": TOY7_V0.copy(), "dt": np.float64(DT), "parent_rows": st.parent_rows.astype(np.int64),
            "dynamic_rows": st.dynamic_rows.astype(np.int64), "diag": d[0], "rhs": s[0], "lowers": lo,
            "uppers": up, "x": x, "V1": o.V.copy()}


def rate_tables():
    V = V_GRID
    na = ch.na_ba2002(V)
    kdr = ch.kdr_ba2002(V)
    cat = ch.cat_hm1992(V)
    cal = ch.cal_is2008(V)
    hcn = ch.hcn_hm1992(V)
    ahp = ch.ahp_de1994(CI_GRID)
    return {"V": V, "Ci": CI_GRID,
            "na_ap": na[0], "na_bp": na[1], "na_aq": na[2], "na_bq": na[3],
            "kdr_ap": kdr[0], "kdr_bp": kdr[1],
            "cat_pinf": cat[0], "cat_ptau": cat[1], "cat_qinf": cat[2], "cat_qtau": cat[3],
            "cal_pinf": cal[0], "cal_ptau": cal[1], "cal_qinf": cal[2], "cal_qtau": cal[3],
            "hcn_pinf": hcn[0], "hcn_ptau": hcn[1],
            "ahp_ap": ahp[0], "ahp_bp": ahp[1],
            "e_ca": ch.nernst(CI_GRID)}


def ca_cell():
    return W.uneven_tree_cell(0, cmin=200, cmax=260)


def main():
    np.savez_compressed(os.path.join(HERE, "c0_layout.npz"), **c0_layout())
    np.savez_compressed(os.path.join(HERE, "c0_trace.npz"), **trace_fixture(W.single_tree_cell(0)))
    np.savez_compressed(os.path.join(HERE, "ca_cell_trace.npz"), **trace_fixture(ca_cell()))
    np.savez_compressed(os.path.join(HERE, "hines7.npz"), **hines7())
    np.savez_compressed(os.path.join(HERE, "rate_tables.npz"), **rate_tables())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
