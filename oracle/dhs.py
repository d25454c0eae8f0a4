"""Oracle restatement of the DHS implicit-Euler voltage solve (FP64).
Test infrastructure only.

Follows ``braincell/quad/_staggered.py``:

- ``_build_dhs_numeric_state`` (:449-513) and ``_edge_current_voltage_delta``
  (:528-539);
- ``comp_triang_raw`` (:641-657): leaf-to-root elimination per level chunk;
- ``comp_backsub_raw`` (:742-760): recursive-doubling back substitution;
- ``_restore_midpoint_voltage`` (:542-543).

A textbook sequential Hines sweep (Hines 1984) is included as an independent
route for the tests.
"""

from __future__ import annotations

import numpy as np


def numeric_state(st, V_cv, linear, const, dt, edge_current_nA=None):
    """_staggered.py:449-513.  Shapes: V_cv/linear/const [batch, C]
    (mV, ms^-1, mV/ms); edge_current_nA [batch, P] in point-id order or None.
    Returns diags, solves [batch, P+1], lowers, uppers [P+1]."""
    V_cv = np.atleast_2d(V_cv)
    linear = np.atleast_2d(linear)
    const = np.atleast_2d(const)
    B = V_cv.shape[0]
    P = st.n_point
    rhs_mid = V_cv + dt * const
    diags = np.broadcast_to(np.concatenate([st.diag * dt, [1.0]])[None, :], (B, P + 1)).copy()
    diags[:, st.dynamic_rows] += 1.0 - dt * linear
    solves = np.zeros((B, P + 1))
    solves[:, st.dynamic_rows] = rhs_mid
    if edge_current_nA is not None:
        cur_by_row = np.atleast_2d(edge_current_nA)[:, st.row_to_point_id]
        # nA / uF = 1e-3 mV/ms
        solves[:, :P] += dt * cur_by_row / st.row_cap_uF * 1e-3
    lowers = np.concatenate([st.lowers * dt, [0.0]])
    uppers = np.concatenate([st.uppers * dt, [0.0]])
    return diags, solves, lowers, uppers


def comp_triang_raw(diags, solves, lowers, uppers, edges, level_offsets):
    """_staggered.py:641-657.  Duplicate parents inside one level are
    accumulated with ``np.add.at`` (a scatter-add)."""
    diags = diags.copy()
    solves = solves.copy()
    for i in range(level_offsets.shape[0] - 1):
        ch = edges[level_offsets[i]:level_offsets[i + 1], 0]
        pa = edges[level_offsets[i]:level_offsets[i + 1], 1]
        mult = uppers[ch] / diags[:, ch]
        np.add.at(diags, (slice(None), pa), -lowers[ch] * mult)
        np.add.at(solves, (slice(None), pa), -solves[:, ch] * mult)
    return diags, solves


def comp_backsub_raw(diags, solves, lowers, backsub):
    """_staggered.py:742-760 (lowers[0] := 0, then pointer jumping)."""
    lowers = lowers.copy()
    lowers[0] = 0.0
    le = -lowers[None, :] / diags
    se = solves / diags
    for i in range(backsub.shape[0]):
        k = backsub[i]
        se = se + le * se[:, k]
        le = le * le[:, k]
    return se


def dhs_solve(st, V_cv, linear, const, dt, edge_current_nA=None):
    """One implicit-Euler DHS voltage solve; returns V_{n+1} on CVs [batch, C]."""
    d, s, lo, up = numeric_state(st, V_cv, linear, const, dt, edge_current_nA)
    d, s = comp_triang_raw(d, s, lo, up, st.edges, st.level_offsets)
    x = comp_backsub_raw(d, s, lo, st.backsub)
    return x[:, st.dynamic_rows]


def dhs_solve_rows(st, diags, solves, lowers, uppers):
    """Full-row solve (elimination + recursive doubling) of an assembled system."""
    d, s = comp_triang_raw(diags, solves, lowers, uppers, st.edges, st.level_offsets)
    return comp_backsub_raw(d, s, lowers, st.backsub)


def hines_sequential(parent_rows, diag, rhs, lowers, uppers):
    """Textbook Hines (1984) solve of a tree system whose rows are numbered
    so that every parent precedes its children (row 0 is the root).

    Row i: diag[i] x_i + lowers[i] x_parent(i) + sum_children uppers[c] x_c = rhs[i].
    """
    n = diag.shape[0]
    d = diag.astype(np.float64).copy()
    r = rhs.astype(np.float64).copy()
    for i in range(n - 1, 0, -1):
        p = parent_rows[i]
        m = uppers[i] / d[i]
        d[p] -= lowers[i] * m
        r[p] -= r[i] * m
    x = np.empty(n)
    x[0] = r[0] / d[0]
    for i in range(1, n):
        x[i] = (r[i] - lowers[i] * x[parent_rows[i]]) / d[i]
    return x


def assemble_dense(st, diags_row, lowers_row, uppers_row):
    """Dense [P, P] matrix of a node-tree system (for brute-force checks)."""
    P = st.n_point
    M = np.diag(diags_row[:P].astype(np.float64))
    for r in range(P):
        p = st.parent_rows[r]
        if p >= 0:
            M[r, p] = lowers_row[r]
            M[p, r] = uppers_row[r]
    return M
