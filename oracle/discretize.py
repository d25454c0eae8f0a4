"""Oracle restatement of CV geometry, node tree, DHS scheduling and the DHS
static source (FP64, NumPy).  Test infrastructure only.

Follows, function by function:

- ``braincell/_discretization/policy.py:84-122, 398-399`` (CVPerBranch bounds)
- ``braincell/_discretization/geometry.py:214-434, 614-736`` (frusta, lateral
  area, axial factors, midpoint split)
- ``braincell/_discretization/base.py:530-559`` (``_assemble_cv``: r_axial =
  Ra * factor)
- ``braincell/_discretization/node_build.py:58-284`` (node tree)
- ``braincell/_compute/scheduling.py:45-229`` (peel levels, row order, groups)
- ``braincell/quad/_staggered.py:323-390, 546-612, 726-739`` (static source)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS_PARAM = 1e-9
EPS_LEN_UM = 1e-6


@dataclass
class Frustum:
    prox: float
    dist: float
    length_um: float
    r_prox_um: float
    r_dist_um: float


@dataclass
class OCV:
    id: int
    branch_id: int
    prox: float
    dist: float
    length_um: float
    area_um2: float
    r_axial_prox: float  # ohm
    r_axial_dist: float  # ohm
    cm: float  # uF/cm^2


def uniform_bounds(n_cv: int):
    """policy.py:398-399."""
    return tuple((float(k) / float(n_cv), float(k + 1) / float(n_cv)) for k in range(n_cv))


def build_frusta(branch, prox: float, dist: float):
    """geometry.py:214-321 (clip a branch to [prox, dist] as frustum pieces)."""
    prox_f = max(0.0, min(1.0, float(prox)))
    dist_f = max(0.0, min(1.0, float(dist)))
    lengths = np.asarray(branch.lengths, dtype=float)
    rp = np.asarray(branch.radii_proximal, dtype=float)
    rd = np.asarray(branch.radii_distal, dtype=float)
    total = float(np.sum(lengths))
    start_um = prox_f * total
    end_um = dist_f * total
    seg_starts = np.concatenate(([0.0], np.cumsum(lengths)[:-1]))
    seg_ends = seg_starts + lengths
    out = []
    for i, seg_len in enumerate(lengths):
        s0 = float(seg_starts[i])
        s1 = float(seg_ends[i])
        if seg_len <= EPS_LEN_UM:
            r0 = float(rp[i])
            r1 = float(rd[i])
            if np.isclose(r0, r1):
                continue
            # geometry.py:200-211 ownership of a zero-length radius jump
            pos = s0
            if pos < start_um - EPS_LEN_UM or pos > end_um + EPS_LEN_UM:
                continue
            owns = (start_um <= EPS_LEN_UM) if pos <= start_um + EPS_LEN_UM else (pos <= end_um + EPS_LEN_UM)
            if not owns:
                continue
            xj = max(prox_f, min(dist_f, s0 / total))
            out.append(Frustum(xj, xj, 0.0, r0, r1))
            continue
        left = max(s0, start_um)
        right = min(s1, end_um)
        if right - left <= EPS_LEN_UM:
            continue
        t0 = (left - s0) / float(seg_len)
        t1 = (right - s0) / float(seg_len)
        r0 = float(rp[i]) + (float(rd[i]) - float(rp[i])) * t0
        r1 = float(rp[i]) + (float(rd[i]) - float(rp[i])) * t1
        x0 = max(prox_f, min(dist_f, left / total))
        x1 = max(prox_f, min(dist_f, right / total))
        out.append(Frustum(x0, x1, float(right - left), r0, r1))
    if not out:
        raise ValueError("CV produced no frusta")
    return out


def lateral_area_um2(frusta) -> float:
    """geometry.py:330-336: sum of pi (r0 + r1) * slant."""
    total = 0.0
    for p in frusta:
        slant = float(np.sqrt(p.length_um ** 2 + (p.r_dist_um - p.r_prox_um) ** 2))
        total += float(np.pi) * (p.r_prox_um + p.r_dist_um) * slant
    return total


def axial_factor_per_cm(frusta) -> float:
    """geometry.py:339-349: sum of L / (pi r0 r1) in cm^-1."""
    f = 0.0
    for p in frusta:
        f += (p.length_um * 1e-4) / (float(np.pi) * (p.r_prox_um * 1e-4) * (p.r_dist_um * 1e-4))
    return f


def split_frusta(frusta, x: float):
    """geometry.py:382-434: split pieces at normalised coordinate x."""
    left, right = [], []
    for p in frusta:
        if p.dist <= x + EPS_PARAM:
            left.append(p)
            continue
        if p.prox >= x - EPS_PARAM:
            right.append(p)
            continue
        span = p.dist - p.prox
        if span <= EPS_PARAM:
            continue
        ratio = max(0.0, min(1.0, (x - p.prox) / span))
        ll = p.length_um * ratio
        lr = p.length_um - ll
        rm = p.r_prox_um + (p.r_dist_um - p.r_prox_um) * ratio
        if ll > EPS_LEN_UM:
            left.append(Frustum(p.prox, x, ll, p.r_prox_um, rm))
        if lr > EPS_LEN_UM:
            right.append(Frustum(x, p.dist, lr, rm, p.r_dist_um))
    return left, right


def build_cvs(morph, cv_per_branch, cm=1.0, ra=100.0):
    """geometry.py:614-736 + base.py:530-559 for a CVPerBranch policy.

    ``cv_per_branch`` may be an int or a per-branch sequence; ``cm`` (uF/cm^2)
    and ``ra`` (ohm cm) may be scalars or per-branch sequences.
    """
    nb = len(morph.branches)
    ncv = [int(cv_per_branch)] * nb if np.isscalar(cv_per_branch) else [int(c) for c in cv_per_branch]
    cm_b = [float(cm)] * nb if np.isscalar(cm) else [float(c) for c in cm]
    ra_b = [float(ra)] * nb if np.isscalar(ra) else [float(r) for r in ra]
    cvs = []
    branch_cv_ids = []
    cid = 0
    for b, br in enumerate(morph.branches):
        ids = []
        for prox, dist in uniform_bounds(ncv[b]):
            fr = build_frusta(br, prox, dist)
            length = sum(p.length_um for p in fr)
            area = lateral_area_um2(fr)
            mid = 0.5 * (prox + dist)
            left, right = split_frusta(fr, mid)
            fprox = axial_factor_per_cm(left) if left else 0.0
            fdist = axial_factor_per_cm(right) if right else 0.0
            cvs.append(OCV(cid, b, prox, dist, length, area, ra_b[b] * fprox, ra_b[b] * fdist, cm_b[b]))
            ids.append(cid)
            cid += 1
        branch_cv_ids.append(tuple(ids))
    return cvs, tuple(branch_cv_ids)


def locate_branch_cv_by_x(ids, cvs, x, eps=EPS_PARAM):
    """node_build.py:396-433 (half-open CV ownership)."""
    if x <= 0.0 + eps:
        return ids[0]
    if x >= 1.0 - eps:
        return ids[-1]
    for cid in ids:
        cv = cvs[cid]
        if cv.prox - eps <= x < cv.dist - eps:
            return cid
    raise ValueError("x lies in no CV interval")


@dataclass
class NodeTree:
    n_nodes: int
    kind: np.ndarray  # 1 = midpoint, 0 = boundary
    edges: list  # (parent_node, child_node, [(cv_id, half), ...]) in creation order
    cv_to_mid_node: np.ndarray
    branch_endpoint_node: np.ndarray  # [n_branch, 2] for x = 0 and x = 1


def build_node_tree(morph, cvs, branch_cv_ids) -> NodeTree:
    """node_build.py:58-284: root endpoint, then per branch (insertion order)
    its midpoints and its terminal endpoint; edges carry CV half roles."""
    roles = []  # per node: set of (cv_id, position)
    cv_to_mid = np.full(len(cvs), -1, dtype=np.int64)
    endpoint = {}
    edge_roles = {}
    edge_order = []

    def new_node(cv_id, pos):
        roles.append({(cv_id, pos)})
        return len(roles) - 1

    def add_edge_role(p, c, cv_id, half):
        key = (p, c)
        if key not in edge_roles:
            edge_roles[key] = []
            edge_order.append(key)
        if (cv_id, half) not in edge_roles[key]:
            edge_roles[key].append((cv_id, half))

    root = new_node(branch_cv_ids[0][0], "prox")
    endpoint[(0, 0.0)] = root
    for b in range(len(morph.branches)):
        ids = branch_cv_ids[b]
        if morph.parent[b] < 0:
            attach = root
            ax = 0.0
            ordered = ids
        else:
            pb = morph.parent[b]
            px = float(morph.parent_x[b])
            if px <= EPS_PARAM:
                attach = endpoint[(pb, 0.0)]
            elif px >= 1.0 - EPS_PARAM:
                attach = endpoint[(pb, 1.0)]
            else:
                attach = int(cv_to_mid[locate_branch_cv_by_x(branch_cv_ids[pb], cvs, px)])
            ax = float(morph.child_x[b])
            ordered = ids if ax <= EPS_PARAM else tuple(reversed(ids))
        entry = "prox" if ax <= EPS_PARAM else "dist"
        exitp = "dist" if ax <= EPS_PARAM else "prox"
        roles[attach].add((ordered[0], entry))
        endpoint[(b, float(ax))] = attach
        for cid in ordered:
            cv_to_mid[cid] = new_node(cid, "mid")
        term = new_node(ordered[-1], exitp)
        endpoint[(b, float(1.0 - ax))] = term
        for k, cid in enumerate(ordered):
            m = int(cv_to_mid[cid])
            p = attach if k == 0 else int(cv_to_mid[ordered[k - 1]])
            c = term if k == len(ordered) - 1 else int(cv_to_mid[ordered[k + 1]])
            add_edge_role(p, m, cid, entry)
            add_edge_role(m, c, cid, exitp)
    n = len(roles)
    kind = np.asarray([1 if all(pos == "mid" for _, pos in r) else 0 for r in roles], dtype=np.int64)
    edges = [(p, c, sorted(edge_roles[(p, c)])) for (p, c) in edge_order]
    bep = np.full((len(morph.branches), 2), -1, dtype=np.int64)
    for b in range(len(morph.branches)):
        bep[b, 0] = endpoint[(b, 0.0)]
        bep[b, 1] = endpoint[(b, 1.0)]
    return NodeTree(n, kind, edges, cv_to_mid, bep)


def edge_conductance(roles, cvs) -> float:
    """_staggered.py:546-576: same-branch halves in series, otherwise parallel (S)."""
    res = []
    branches = set()
    for cid, half in roles:
        cv = cvs[cid]
        branches.add(cv.branch_id)
        res.append(cv.r_axial_prox if half == "prox" else cv.r_axial_dist)
    if len(res) > 1 and len(branches) == 1:
        tot = None
        for r in res:
            tot = r if tot is None else tot + r
        return 1.0 / tot
    g = None
    for r in res:
        v = 1.0 / r
        g = v if g is None else g + v
    return g


def node_parent_children(tree: NodeTree):
    """scheduling.py:216-229."""
    parent = np.full(tree.n_nodes, -1, dtype=np.int64)
    children = [[] for _ in range(tree.n_nodes)]
    for p, c, _ in tree.edges:
        if parent[c] != -1:
            raise ValueError(f"Node {c} already has parent {parent[c]}.")
        parent[c] = p
        children[p].append(c)
    return parent, children


def compute_peel_levels(node_parent, node_children) -> np.ndarray:
    """scheduling.py:120-159: distance to farthest leaf, BFS from the leaves."""
    n = len(node_parent)
    levels = np.full(n, -1, dtype=np.int64)
    remaining = np.asarray([len(c) for c in node_children], dtype=np.int64)
    frontier = [i for i in range(n) if remaining[i] == 0]
    for i in frontier:
        levels[i] = 0
    cur = 0
    while cur < len(frontier):
        pid = frontier[cur]
        cur += 1
        par = int(node_parent[pid])
        if par < 0:
            continue
        cand = int(levels[pid]) + 1
        if levels[par] < cand:
            levels[par] = cand
        remaining[par] -= 1
        if remaining[par] == 0:
            frontier.append(par)
    if (levels < 0).any():
        raise ValueError("compute_peel_levels: cycle detected or node unreachable from any leaf.")
    return levels


@dataclass
class Scheduling:
    row_to_point_id: np.ndarray
    point_id_to_row: np.ndarray
    groups: tuple
    parent_rows: np.ndarray


def build_scheduling(tree: NodeTree, max_group_size: int = 256) -> Scheduling:
    """scheduling.py:45-105, 162-213 (Cell default max_group_size=256,
    ``_multi_compartment/cell.py:1282-1297``)."""
    parent, children = node_parent_children(tree)
    peel = compute_peel_levels(parent, children)
    order = np.lexsort((np.arange(tree.n_nodes), -peel)).astype(np.int64)
    p2r = np.empty(tree.n_nodes, dtype=np.int64)
    p2r[order] = np.arange(tree.n_nodes)
    parent_rows = np.full(tree.n_nodes, -1, dtype=np.int64)
    for row, pid in enumerate(order.tolist()):
        if parent[pid] >= 0:
            parent_rows[row] = p2r[parent[pid]]
    peels_o = peel[order]
    starts = [0] + [i for i in range(1, len(order)) if peels_o[i] != peels_o[i - 1]] + [len(order)]
    groups = []
    for a, b in zip(starts[:-1], starts[1:]):
        for cs in range(a, b, max_group_size):
            ce = min(b, cs + max_group_size)
            groups.append(p2r[order[cs:ce]])
    return Scheduling(order, p2r, tuple(groups), parent_rows)


def build_backsub_indices(parent_lookup, n_nodes: int) -> np.ndarray:
    """_staggered.py:726-739: recursive-doubling ancestor jump table."""
    parent_lookup = np.asarray(parent_lookup, dtype=np.int64)
    out = []
    old, new = 0, 1
    k = np.arange(n_nodes + 1, dtype=np.int64)
    while new <= max(1, n_nodes):
        for _ in range(new - old):
            k = parent_lookup[k]
        old = new
        new = 2 * new
        out.append(k)
    return np.asarray(out, dtype=np.int64)


def build_edge_order(sched: Scheduling):
    """_staggered.py:596-612: leaf-to-root elimination edges per group."""
    pairs = []
    sizes = []
    for g in reversed(sched.groups):
        lev = [[int(r), int(sched.parent_rows[r])] for r in g.tolist() if sched.parent_rows[r] >= 0]
        if lev:
            pairs.extend(lev)
            sizes.append(len(lev))
    if pairs:
        return np.asarray(pairs, dtype=np.int64).reshape(-1, 2), np.asarray(sizes, dtype=np.int64)
    return np.empty((0, 2), dtype=np.int64), np.empty((0,), dtype=np.int64)


@dataclass
class DHSStatic:
    n_point: int
    dynamic_rows: np.ndarray  # CV -> row
    row_to_point_id: np.ndarray
    row_cap_uF: np.ndarray
    diag: np.ndarray  # ms^-1
    lowers: np.ndarray
    uppers: np.ndarray
    axial: np.ndarray  # dense [P, P] ms^-1
    edges: np.ndarray
    level_offsets: np.ndarray
    backsub: np.ndarray
    parent_rows: np.ndarray


def build_dhs_static(cvs, tree: NodeTree, sched: Scheduling) -> DHSStatic:
    """_staggered.py:323-390, 579-593: dense axial operator in ms^-1.

    Row capacitance: CV midpoints use area * cm (uF); boundary rows use a
    dummy 1 uF.  Coefficient = g[S] / C[uF] = 1e3 * g / C in ms^-1.
    """
    n = tree.n_nodes
    p2r = sched.point_id_to_row
    dyn = np.asarray([int(p2r[tree.cv_to_mid_node[c]]) for c in range(len(cvs))], dtype=np.int64)
    cap = np.ones(n, dtype=np.float64)
    for c, r in enumerate(dyn.tolist()):
        cap[r] = cvs[c].area_um2 * 1e-8 * cvs[c].cm  # um^2 -> cm^2, times uF/cm^2
    A = np.zeros((n, n), dtype=np.float64)
    for p, c, roles in tree.edges:
        pr, cr = int(p2r[p]), int(p2r[c])
        g = edge_conductance(roles, cvs)
        pc = g / cap[pr] * 1e-3 * 1e6  # S/uF = 1e6 s^-1 = 1e3 ms^-1
        cc = g / cap[cr] * 1e-3 * 1e6
        A[pr, pr] += pc
        A[pr, cr] -= pc
        A[cr, cr] += cc
        A[cr, pr] -= cc
    diag = np.asarray(np.diag(A), dtype=np.float64).copy()
    lowers = np.zeros(n)
    uppers = np.zeros(n)
    for row, pr in enumerate(sched.parent_rows.tolist()):
        if pr < 0:
            continue
        lowers[row] = A[row, pr]
        uppers[row] = A[pr, row]
    lookup = np.empty(n + 1, dtype=np.int64)
    lookup[:n] = np.where(sched.parent_rows >= 0, sched.parent_rows, n)
    lookup[n] = n
    edges, sizes = build_edge_order(sched)
    back = build_backsub_indices(lookup, n)
    offs = np.cumsum(np.insert(sizes, 0, 0)).astype(np.int64)
    return DHSStatic(n, dyn, sched.row_to_point_id, cap, diag, lowers, uppers, A, edges, offs, back, sched.parent_rows)


def schur_cv_operator(st: DHSStatic) -> np.ndarray:
    """_staggered.py:393-413: reduce the mixed node-tree operator onto CV rows."""
    dyn = st.dynamic_rows
    alg = np.asarray([r for r in range(st.n_point) if r not in set(dyn.tolist())], dtype=np.int64)
    A = st.axial
    if alg.size == 0:
        return A[np.ix_(dyn, dyn)]
    return A[np.ix_(dyn, dyn)] - A[np.ix_(dyn, alg)] @ np.linalg.solve(A[np.ix_(alg, alg)], A[np.ix_(alg, dyn)])


def discretize(morph, cv_per_branch=4, cm=1.0, ra=100.0, max_group_size=256):
    """Convenience: CVs, node tree, scheduling and DHS static source."""
    cvs, bids = build_cvs(morph, cv_per_branch, cm, ra)
    tree = build_node_tree(morph, cvs, bids)
    sched = build_scheduling(tree, max_group_size)
    st = build_dhs_static(cvs, tree, sched)
    return cvs, bids, tree, sched, st
