"""Flat, structure-of-arrays layout of one or many cells for the HIP kernels.

This is the re-laid-out output of braincell's ``_discretization`` +
``_compute.scheduling`` + DHS static source:

- CV geometry per ``CVPerBranch`` (``_discretization/geometry.py:614-736``,
  ``base.py:530-559``): lateral frustum areas and the two half-CV axial
  resistances;
- the node tree (``_discretization/node_build.py:58-284``): root endpoint,
  then per branch its CV midpoints and its terminal endpoint;
- the DHS row order (``_compute/scheduling.py:120-171``): peel level
  descending, then node id, i.e. exactly the reference ``row_to_point_id``;
- the axial coefficients in ms^-1 (``quad/_staggered.py:323-390, 546-593``):
  ``diag_axial``, ``lowers[row] = A[row, parent]``, ``uppers[row] = A[parent,
  row]``.

On top of the reference rows it derives the *chain schedule* the Hines
kernel runs: maximal runs of rows whose parent has exactly one child are
eliminated by one thread in registers, and chains are grouped into levels
(a chain's level is one more than the deepest chain hanging below it).

Everything here is host-side NumPy, computed once per cell.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS_PARAM = 1e-9
EPS_LEN_UM = 1e-6

__all__ = ["CellGeometry", "discretize_cell", "build_chains", "build_chains_batch", "Chains"]


@dataclass
class CellGeometry:
    """One cell's discretisation and DHS rows (FP64, reference order)."""

    n_cv: int
    n_point: int
    cv_branch: np.ndarray  # [C] branch index of each CV
    cv_area_um2: np.ndarray  # [C]
    cv_cm: np.ndarray  # [C] uF/cm^2
    r_axial_prox: np.ndarray  # [C] ohm
    r_axial_dist: np.ndarray  # [C] ohm
    node_parent: np.ndarray  # [P] parent node id (-1 root)
    node_is_mid: np.ndarray  # [P] bool
    cv_to_node: np.ndarray  # [C]
    branch_endpoint: np.ndarray  # [B, 2] node id at x=0 / x=1
    peel: np.ndarray  # [P] per node
    row_to_point_id: np.ndarray  # [P]
    point_id_to_row: np.ndarray  # [P]
    parent_rows: np.ndarray  # [P] (-1 root)
    dynamic_rows: np.ndarray  # [C] CV -> row
    row_cap_uF: np.ndarray  # [P]
    diag_axial: np.ndarray  # [P] ms^-1
    lowers: np.ndarray  # [P] ms^-1
    uppers: np.ndarray  # [P] ms^-1
    row_is_mid: np.ndarray  # [P] bool (row indexed)

    def locate(self, branch: int, x: float, n_cv_branch: np.ndarray):
        """Resolve a placed point to ("mid", cv) or ("edge", row)
        (``mechanism.py:506-511``, ``node_build.py:232-243, 396-433``)."""
        if x <= 0.0 + EPS_PARAM:
            return "edge", int(self.point_id_to_row[self.branch_endpoint[branch, 0]])
        if x >= 1.0 - EPS_PARAM:
            return "edge", int(self.point_id_to_row[self.branch_endpoint[branch, 1]])
        first = int(np.sum(n_cv_branch[:branch]))
        n = int(n_cv_branch[branch])
        for k in range(n):
            prox = float(k) / float(n)
            dist = float(k + 1) / float(n)
            if prox - EPS_PARAM <= x < dist - EPS_PARAM:
                return "mid", first + k
        raise ValueError(f"x={x} lies in no CV of branch {branch}")


def _cv_geometry_vectorized(L, r0, r1, k, n, ra):
    """Closed form of the frustum pipeline for CV ``k`` of ``n`` uniform CVs on
    one-segment branches (geometry.py:214-434 specialised); all arrays are
    per CV."""
    prox = k / n
    dist = (k + 1.0) / n
    left = prox * L
    right = dist * L
    t0 = left / L
    t1 = right / L
    ra_ = r0 + (r1 - r0) * t0
    rb_ = r0 + (r1 - r0) * t1
    length = right - left
    area = np.pi * (ra_ + rb_) * np.sqrt(length ** 2 + (rb_ - ra_) ** 2)
    mid = 0.5 * (prox + dist)
    span = dist - prox
    ratio = np.clip((mid - prox) / span, 0.0, 1.0)
    ll = length * ratio
    lr = length - ll
    rm = ra_ + (rb_ - ra_) * ratio
    fprox = (ll * 1e-4) / (np.pi * (ra_ * 1e-4) * (rm * 1e-4))
    fdist = (lr * 1e-4) / (np.pi * (rm * 1e-4) * (rb_ * 1e-4))
    return area, ra * fprox, ra * fdist


def _frusta(lengths, rp, rd, prox, dist):
    total = float(np.sum(lengths))
    s_um, e_um = prox * total, dist * total
    starts = np.concatenate(([0.0], np.cumsum(lengths)[:-1]))
    out = []
    for i, sl in enumerate(lengths):
        s0 = float(starts[i])
        s1 = s0 + float(sl)
        if sl <= EPS_LEN_UM:
            if np.isclose(rp[i], rd[i]):
                continue
            if s0 < s_um - EPS_LEN_UM or s0 > e_um + EPS_LEN_UM:
                continue
            owns = (s_um <= EPS_LEN_UM) if s0 <= s_um + EPS_LEN_UM else (s0 <= e_um + EPS_LEN_UM)
            if owns:
                xj = max(prox, min(dist, s0 / total))
                out.append((xj, xj, 0.0, float(rp[i]), float(rd[i])))
            continue
        lo, hi = max(s0, s_um), min(s1, e_um)
        if hi - lo <= EPS_LEN_UM:
            continue
        t0, t1 = (lo - s0) / float(sl), (hi - s0) / float(sl)
        r0 = float(rp[i]) + (float(rd[i]) - float(rp[i])) * t0
        r1 = float(rp[i]) + (float(rd[i]) - float(rp[i])) * t1
        out.append((max(prox, min(dist, lo / total)), max(prox, min(dist, hi / total)), hi - lo, r0, r1))
    return out


def _cv_geometry_general(br, n, ra):
    """General multi-segment frustum pipeline (geometry.py:214-434)."""
    area = np.zeros(n)
    rprox = np.zeros(n)
    rdist = np.zeros(n)
    for k in range(n):
        prox, dist = float(k) / float(n), float(k + 1) / float(n)
        fr = _frusta(br.lengths, br.radii_proximal, br.radii_distal, prox, dist)
        a = 0.0
        for (_, _, ln, r0, r1) in fr:
            a += np.pi * (r0 + r1) * np.sqrt(ln ** 2 + (r1 - r0) ** 2)
        area[k] = a
        mid = 0.5 * (prox + dist)
        left, right = [], []
        for (p0, p1, ln, r0, r1) in fr:
            if p1 <= mid + EPS_PARAM:
                left.append((ln, r0, r1))
                continue
            if p0 >= mid - EPS_PARAM:
                right.append((ln, r0, r1))
                continue
            span = p1 - p0
            if span <= EPS_PARAM:
                continue
            ratio = max(0.0, min(1.0, (mid - p0) / span))
            ll = ln * ratio
            lr = ln - ll
            rm = r0 + (r1 - r0) * ratio
            if ll > EPS_LEN_UM:
                left.append((ll, r0, rm))
            if lr > EPS_LEN_UM:
                right.append((lr, rm, r1))
        f = lambda pcs: sum((ln * 1e-4) / (np.pi * (r0 * 1e-4) * (r1 * 1e-4)) for ln, r0, r1 in pcs)
        rprox[k] = ra * f(left)
        rdist[k] = ra * f(right)
    return area, rprox, rdist


def _owning_cv(x: float, n: int) -> int:
    """CV k of n uniform CVs owns [k/n, (k+1)/n) with the reference's
    tolerance (``node_build.py:396-433``)."""
    if x <= EPS_PARAM:
        return 0
    if x >= 1.0 - EPS_PARAM:
        return n - 1
    for k in range(n):
        if float(k) / n - EPS_PARAM <= x < float(k + 1) / n - EPS_PARAM:
            return k
    raise ValueError(f"x={x} lies in no CV")


def discretize_cell(morph, n_cv_branch, cm_branch, ra_branch) -> CellGeometry:
    """Discretise one cell and build its reference-ordered DHS rows."""
    nb = morph.n_branches
    ncv = np.asarray(n_cv_branch, dtype=np.int64)
    if ncv.shape != (nb,) or np.any(ncv <= 0):
        raise ValueError("n_cv_branch must give a positive CV count per branch")
    for b in range(1, nb):
        if morph.parent[b] >= b:
            raise ValueError("branches must be inserted after their parent")
    C = int(ncv.sum())
    cv_first = np.concatenate(([0], np.cumsum(ncv)[:-1]))
    cv_branch = np.repeat(np.arange(nb), ncv)
    kk = np.arange(C) - cv_first[cv_branch]
    cm = np.asarray(cm_branch, dtype=np.float64)[cv_branch]
    ra_cv = np.asarray(ra_branch, dtype=np.float64)[cv_branch]
    # single-segment branches: the frustum pipeline in closed form, all CVs at once
    nseg = np.asarray([br.lengths.shape[0] for br in morph.branches])
    Lb = np.asarray([float(br.lengths[0]) for br in morph.branches])
    r0b = np.asarray([float(br.radii_proximal[0]) for br in morph.branches])
    r1b = np.asarray([float(br.radii_distal[0]) for br in morph.branches])
    area, rprox, rdist = _cv_geometry_vectorized(Lb[cv_branch], r0b[cv_branch], r1b[cv_branch],
                                                 kk.astype(np.float64), ncv[cv_branch].astype(np.float64), ra_cv)
    for b in np.nonzero(nseg > 1)[0].tolist():
        s_, n = int(cv_first[b]), int(ncv[b])
        a, rp_, rd_ = _cv_geometry_general(morph.branches[b], n, float(ra_branch[b]))
        area[s_:s_ + n], rprox[s_:s_ + n], rdist[s_:s_ + n] = a, rp_, rd_

    # ---- node tree (node_build.py:58-284).  A child attached inside its
    # parent branch (0 < parent_x < 1) hangs from the midpoint node of the
    # parent CV that owns parent_x under the half-open tiling
    # (node_build.py:396-433); that node stays the CV's dynamic row.
    P = 1 + C + nb
    node_parent = np.full(P, -1, dtype=np.int64)
    node_is_mid = np.zeros(P, dtype=bool)
    node_g = np.zeros(P)  # conductance of the edge to the parent (S)
    cx = np.asarray(morph.child_x, dtype=np.float64)
    px = np.asarray(morph.parent_x, dtype=np.float64)
    par = np.asarray(morph.parent, dtype=np.int64)
    interior = (px > EPS_PARAM) & (px < 1.0 - EPS_PARAM)
    interior[0] = False
    if np.all(cx <= EPS_PARAM) and not np.any(interior):
        base = 1 + np.concatenate(([0], np.cumsum(ncv + 1)[:-1]))
        term = base + ncv
        mid0 = base
        attach = np.zeros(nb, dtype=np.int64)
        if nb > 1:
            if np.all(px[1:] >= 1.0 - EPS_PARAM):
                attach[1:] = term[par[1:]]
            else:
                for b in range(1, nb):
                    attach[b] = term[par[b]] if px[b] >= 1.0 - EPS_PARAM else attach[par[b]]
        cv_to_node = base[cv_branch] + kk
        first = kk == 0
        node_parent[cv_to_node] = np.where(first, attach[cv_branch], cv_to_node - 1)
        prev = np.maximum(np.arange(C) - 1, 0)
        node_g[cv_to_node] = np.where(first, 1.0 / rprox, 1.0 / (rdist[prev] + rprox))
        last_cv = cv_first + ncv - 1
        node_parent[term] = term - 1
        node_g[term] = 1.0 / rdist[last_cv]
        node_is_mid[cv_to_node] = True
        endpoint = np.stack([attach, term], axis=1)
    else:
        cv_to_node = np.empty(C, dtype=np.int64)
        endpoint = np.full((nb, 2), -1, dtype=np.int64)
        mid0 = np.empty(nb, dtype=np.int64)
        attach = np.empty(nb, dtype=np.int64)
        term = np.empty(nb, dtype=np.int64)
        nid = 1
        for b in range(nb):
            n = int(ncv[b])
            ids = np.arange(cv_first[b], cv_first[b] + n)
            if par[b] < 0:
                a_node, ax = 0, 0.0
            else:
                if interior[b]:
                    a_node = int(cv_to_node[cv_first[par[b]] + _owning_cv(float(px[b]), int(ncv[par[b]]))])
                else:
                    a_node = endpoint[par[b], 0] if px[b] <= EPS_PARAM else endpoint[par[b], 1]
                ax = float(cx[b])
            ordered = ids if ax <= EPS_PARAM else ids[::-1]
            r_entry = rprox if ax <= EPS_PARAM else rdist
            r_exit = rdist if ax <= EPS_PARAM else rprox
            side = 0 if ax <= EPS_PARAM else 1
            endpoint[b, side] = a_node
            mids = np.arange(nid, nid + n)
            t_node = nid + n
            nid += n + 1
            endpoint[b, 1 - side] = t_node
            cv_to_node[ordered] = mids
            node_is_mid[mids] = True
            node_parent[mids[0]] = a_node
            node_g[mids[0]] = 1.0 / r_entry[ordered[0]]
            if n > 1:
                node_parent[mids[1:]] = mids[:-1]
                node_g[mids[1:]] = 1.0 / (r_exit[ordered[:-1]] + r_entry[ordered[1:]])
            node_parent[t_node] = mids[-1]
            node_g[t_node] = 1.0 / r_exit[ordered[-1]]
            mid0[b], attach[b], term[b] = mids[0], a_node, t_node

    # ---- peel levels (scheduling.py:120-159) by a branch-level sweep
    peel = np.zeros(P, dtype=np.int64)
    best_child = {}
    for b in range(nb - 1, -1, -1):
        t_node = term[b]
        pt = best_child.get(int(t_node), -1) + 1 if int(t_node) in best_child else 0
        peel[t_node] = pt
        n = int(ncv[b])
        nodes = np.arange(mid0[b], mid0[b] + n)
        if any(int(m) in best_child for m in nodes.tolist()):
            # children attached at this branch's midpoints
            v = pt
            for i in range(n - 1, -1, -1):
                m = int(nodes[i])
                v = max(v, best_child.get(m, -1)) + 1
                peel[m] = v
        else:
            peel[nodes] = pt + (n - np.arange(n))
        a = int(attach[b])
        best_child[a] = max(best_child.get(a, -1), int(peel[mid0[b]]))
    peel[0] = best_child[0] + 1 if 0 in best_child else 0

    order = np.lexsort((np.arange(P), -peel)).astype(np.int64)
    p2r = np.empty(P, dtype=np.int64)
    p2r[order] = np.arange(P)
    par_node = node_parent[order]
    parent_rows = np.where(par_node >= 0, p2r[np.maximum(par_node, 0)], -1)
    dyn = p2r[cv_to_node]

    # ---- axial coefficients (_staggered.py:360-390, 579-593)
    cap = np.ones(P)
    cap[dyn] = area * 1e-8 * cm
    child_rows = np.nonzero(parent_rows >= 0)[0]
    g_rows = node_g[order]
    # edges in creation order = child node id order (each node has one parent edge)
    child_nodes = np.nonzero(node_parent >= 0)[0]
    crow = p2r[child_nodes]
    prow = p2r[node_parent[child_nodes]]
    g = node_g[child_nodes]
    pc = g / cap[prow] * 1e3
    cc = g / cap[crow] * 1e3
    diag = np.zeros(P)
    # the dense assembly adds parent then child per edge, in edge order
    inter = np.empty(2 * len(child_nodes), dtype=np.int64)
    vals = np.empty(2 * len(child_nodes))
    inter[0::2], inter[1::2] = prow, crow
    vals[0::2], vals[1::2] = pc, cc
    np.add.at(diag, inter, vals)
    lowers = np.zeros(P)
    uppers = np.zeros(P)
    lowers[crow] = -cc
    uppers[crow] = -pc
    del child_rows, g_rows
    return CellGeometry(
        n_cv=C, n_point=P, cv_branch=cv_branch, cv_area_um2=area, cv_cm=cm, r_axial_prox=rprox,
        r_axial_dist=rdist, node_parent=node_parent, node_is_mid=node_is_mid, cv_to_node=cv_to_node,
        branch_endpoint=endpoint, peel=peel, row_to_point_id=order, point_id_to_row=p2r,
        parent_rows=parent_rows, dynamic_rows=dyn, row_cap_uF=cap, diag_axial=diag, lowers=lowers,
        uppers=uppers, row_is_mid=node_is_mid[order])


@dataclass
class Chains:
    """Chain schedule over one cell's (or one pack's) rows.

    ``rows[off[k]:off[k+1]]`` lists chain ``k`` bottom-to-top; ``top_parent[k]``
    is the row the chain's top hangs from (-1 for a root); the chain's bottom
    gathers the contributions of chains ``kids[kid_off[k]:kid_off[k+1]]``.
    Chains are sorted by level; ``level_off`` delimits levels.
    """

    off: np.ndarray
    rows: np.ndarray
    top_parent: np.ndarray
    kid_off: np.ndarray
    kids: np.ndarray
    level_off: np.ndarray
    level: np.ndarray

    @property
    def n_chains(self) -> int:
        return int(self.off.shape[0] - 1)

    @property
    def n_levels(self) -> int:
        return int(self.level_off.shape[0] - 1)


def build_chains(parent_rows: np.ndarray, max_len: int = None) -> Chains:
    """Split a row forest (parents given per row; -1 for roots) into chains:
    maximal runs in which every row is the only child of the next, cut from
    the bottom into pieces of at most ``max_len`` rows (a piece's top then
    hangs from the bottom row of the piece above)."""
    return build_chains_batch([parent_rows], max_len)[0]


def build_chains_batch(parent_rows_list, max_len: int = None) -> list:
    """``build_chains`` for many row forests at once (vectorised over all of
    their rows).  Chains of each forest are sorted by (level, bottom row),
    level = height in the chain tree (leaves 0); kids are listed in
    increasing chain index."""
    cap = int(max_len) if max_len else None
    if cap is not None and cap < 1:
        raise ValueError("max_len must be >= 1")
    sizes = np.asarray([len(p) for p in parent_rows_list], dtype=np.int64)
    n_f = len(sizes)
    base = np.concatenate(([0], np.cumsum(sizes)))
    N = int(base[-1])
    if N == 0:
        raise ValueError("empty row forest")
    loc = np.concatenate([np.asarray(p, dtype=np.int64) for p in parent_rows_list])
    forest = np.repeat(np.arange(n_f), sizes)
    par = np.where(loc >= 0, loc + base[forest], -1)
    has_p = par >= 0
    nchild = np.bincount(par[has_p], minlength=N)
    # a row continues its parent's chain when it is that parent's only child
    up = np.where(has_p, par, 0)
    cont = has_p & (nchild[up] == 1)
    prev = np.full(N, -1, dtype=np.int64)
    prev[par[cont]] = np.nonzero(cont)[0]
    # list ranking: distance of every row to the bottom of its maximal chain
    low = np.where(prev >= 0, prev, np.arange(N))
    dist = (prev >= 0).astype(np.int64)
    while True:
        nxt = low[low]
        if np.array_equal(nxt, low):
            break
        dist = dist + dist[low] * (low != np.arange(N))
        low = nxt
    # (the loop adds dist[low] while low has not reached the bottom)
    bottom = low
    piece = dist // cap if cap else np.zeros(N, dtype=np.int64)
    # rows in chain order: by (bottom, distance); pieces are contiguous runs
    order = np.lexsort((dist, bottom))
    b_o, p_o = bottom[order], piece[order]
    start = np.ones(N, dtype=bool)
    start[1:] = (b_o[1:] != b_o[:-1]) | (p_o[1:] != p_o[:-1])
    chain_of_pos = np.cumsum(start) - 1
    n_ch = int(chain_of_pos[-1]) + 1
    chain_of_row = np.empty(N, dtype=np.int64)
    chain_of_row[order] = chain_of_pos
    first = np.nonzero(start)[0]
    last = np.concatenate((first[1:], [N])) - 1
    bot_row = order[first]
    top_row = order[last]
    top_parent = par[top_row]
    tp_ok = top_parent >= 0
    parent_chain = np.where(tp_ok, chain_of_row[np.maximum(top_parent, 0)], -1)
    # levels: height in the chain tree, by relaxation over the kid edges
    kid_idx = np.nonzero(parent_chain >= 0)[0]
    kpar = parent_chain[kid_idx]
    srt = np.argsort(kpar, kind="stable")
    kid_idx, kpar = kid_idx[srt], kpar[srt]
    level = np.zeros(n_ch, dtype=np.int64)
    if kid_idx.size:
        grp = np.nonzero(np.concatenate(([True], kpar[1:] != kpar[:-1])))[0]
        gpar = kpar[grp]
        while True:
            cand = np.maximum.reduceat(level[kid_idx] + 1, grp)
            new = level.copy()
            new[gpar] = np.maximum(new[gpar], cand)
            if np.array_equal(new, level):
                break
            level = new
    ch_forest = forest[bot_row]
    out = []
    ch_base = np.concatenate(([0], np.cumsum(np.bincount(ch_forest, minlength=n_f))))
    lens_all = last - first + 1
    for f in range(n_f):
        c0, c1 = int(ch_base[f]), int(ch_base[f + 1])
        b0 = base[f]
        bottoms = bot_row[c0:c1] - b0
        lv = level[c0:c1]
        perm = np.lexsort((bottoms, lv))
        nch = c1 - c0
        inv = np.empty(nch, dtype=np.int64)
        inv[perm] = np.arange(nch)
        lens = lens_all[c0:c1][perm]
        off = np.concatenate(([0], np.cumsum(lens)))
        # rows of the forest's chains in the new chain order
        pos0 = first[c0:c1][perm]
        idx = np.repeat(pos0 - off[:-1], lens) + np.arange(int(off[-1]))
        rows = order[idx] - b0
        tp = np.where(tp_ok[c0:c1], top_parent[c0:c1] - b0, -1)[perm]
        pch = parent_chain[c0:c1][perm]
        pch_new = np.where(pch >= 0, inv[np.maximum(pch - c0, 0)], -1)
        kid_cnt = np.bincount(pch_new[pch_new >= 0], minlength=nch)
        kid_off = np.concatenate(([0], np.cumsum(kid_cnt)))
        kids = np.argsort(np.where(pch_new >= 0, pch_new, nch), kind="stable")[: int(kid_cnt.sum())]
        lvs = lv[perm]
        nlev = int(lvs.max()) + 1
        level_off = np.searchsorted(lvs, np.arange(nlev + 1), side="left")
        out.append(Chains(off=off, rows=rows, top_parent=tp, kid_off=kid_off, kids=kids,
                          level_off=level_off.astype(np.int64), level=lvs))
    return out


