"""Product layout vs the oracle's restatement, chain invariants, and the CPU
transcription of the kernel's Hines phase vs the oracle DHS (CPU only)."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from braincell_amd.layout import build_chains, discretize_cell
from braincell_amd.runtime import StaggerLayout, first_step_at
from oracle import discretize as OD
from oracle.dhs import dhs_solve
from oracle.staggered import StaggeredOracle
from tests.kernel_emulator import bnd_rows, cv_rows, hines_delta, hines_transposed, hines_transposed_composite

CELLS = [lambda: W.single_tree_cell(0)] + [lambda i=i: W.uneven_tree_cell(i, cmin=200, cmax=700) for i in range(4)]


@pytest.mark.parametrize("maker", CELLS)
def test_rows_and_coefficients_match_oracle(maker):
    m = maker()
    g = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    cvs, _, tree, sched, st = OD.discretize(m.morph, 4, 1.0, 100.0)
    assert np.array_equal(g.row_to_point_id, sched.row_to_point_id)  # bit-identical node ordering
    assert np.array_equal(g.parent_rows, sched.parent_rows)
    assert np.array_equal(g.dynamic_rows, st.dynamic_rows)
    for a, b in [(g.diag_axial, st.diag), (g.lowers, st.lowers), (g.uppers, st.uppers), (g.row_cap_uF, st.row_cap_uF)]:
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=0)
    np.testing.assert_allclose(g.cv_area_um2, [c.area_um2 for c in cvs], rtol=1e-14)


def test_multisegment_branch_geometry_matches_oracle():
    from braincell_amd.morph import Branch, Morphology

    soma = Branch.from_lengths(lengths=[10.0, 0.0, 12.0], radii_proximal=[5.0, 5.0, 3.0],
                               radii_distal=[5.0, 3.0, 2.0], type="soma")
    m = Morphology.from_root(soma)
    m.attach(0, Branch.from_lengths(lengths=[30.0, 40.0], radii=[1.0, 0.8, 0.5], type="dendrite"))
    m.attach(0, Branch.from_lengths(lengths=[25.0], radii=[1.2, 0.7], type="dendrite"))
    g = discretize_cell(m, np.asarray([3, 5, 2]), np.ones(3), np.full(3, 120.0))
    cvs, _, _, sched, st = OD.discretize(m, [3, 5, 2], 1.0, 120.0)
    assert np.array_equal(g.row_to_point_id, sched.row_to_point_id)
    np.testing.assert_allclose(g.cv_area_um2, [c.area_um2 for c in cvs], rtol=1e-13)
    np.testing.assert_allclose(g.diag_axial, st.diag, rtol=1e-12)


@pytest.mark.parametrize("max_len", [None, 16, 3, 1])
@pytest.mark.parametrize("maker", CELLS)
def test_chain_invariants(maker, max_len):
    m = maker()
    g = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    ch = build_chains(g.parent_rows, max_len)
    assert sorted(ch.rows.tolist()) == list(range(g.n_point))  # each row in exactly one chain
    full = build_chains(g.parent_rows)
    assert ch.n_chains >= full.n_chains
    if max_len:
        assert np.diff(ch.off).max() <= max_len
    else:  # maximal chains: a chain's top parent has several children (or none)
        nchild = np.bincount(g.parent_rows[g.parent_rows >= 0], minlength=g.n_point)
        assert all(tp < 0 or nchild[tp] != 1 for tp in ch.top_parent.tolist())
    for k in range(ch.n_chains):
        rows = ch.rows[ch.off[k]:ch.off[k + 1]]
        for a, b in zip(rows[:-1], rows[1:]):
            assert g.parent_rows[a] == b  # consecutive rows are child -> parent
        assert ch.top_parent[k] == g.parent_rows[rows[-1]]
        for kid in ch.kids[ch.kid_off[k]:ch.kid_off[k + 1]]:
            assert ch.top_parent[kid] == rows[0]  # kids hang from the bottom row
            assert ch.level[kid] < ch.level[k]
    assert np.all(np.diff(ch.level) >= 0)
    assert ch.level_off[-1] == ch.n_chains
    # every chain of level h > 0 has a kid of level h - 1, so level widths never
    # grow towards the root (the kernel runs the narrow tail in one wavefront)
    assert np.all(np.diff(np.diff(ch.level_off)) <= 0)


def test_pack_level_widths_never_grow():
    models = [W.uneven_tree_cell(i, cmin=200, cmax=600) for i in range(8)]
    L = StaggerLayout(models, rows_per_pack=1500, cv_per_thread=4)
    lvo = np.asarray(L.level_off)
    for pk in L.packs:
        w = np.diff(lvo[pk.lv0:pk.lv0 + pk.n_lv + 1])
        assert np.all(w > 0) and np.all(np.diff(w) <= 0)


@pytest.mark.parametrize("max_chain_rows", [16, 2, None])
def test_packed_hines_emulation_matches_oracle_dhs(max_chain_rows):
    models = [W.uneven_tree_cell(i, cmin=200, cmax=400) for i in range(5)] + [W.single_tree_cell(0)]
    L = StaggerLayout(models, rows_per_pack=1200, cv_per_thread=4, max_chain_rows=max_chain_rows)
    assert L.n_packs < len(models)  # several cells share a pack
    dt = 0.025
    rng = np.random.default_rng(0)
    G = rng.uniform(0, 1, L.n_cv_total)
    V = rng.uniform(-80, 0, L.n_cv_total)
    I = rng.uniform(-5, 5, L.n_cv_total)
    rows, brow = cv_rows(L), bnd_rows(L)
    bd = np.zeros(L.n_rows_total)
    br = np.zeros(L.n_rows_total)
    xv = np.zeros(L.n_rows_total)
    bd[rows] = 1 + dt * G
    br[rows] = dt * I
    xv[rows] = V
    xv[brow] = rng.uniform(-80, 0, len(brow))  # any start value: constraints are re-imposed
    dV = hines_delta(L, dt, bd, br, xv)
    for i, m in enumerate(models):
        st = StaggeredOracle(m).st
        sl = L.cell_slice(i)
        Vn = dhs_solve(st, V[sl][None], (-G[sl])[None], (I[sl] + G[sl] * V[sl])[None], dt)[0]
        np.testing.assert_allclose(V[sl] + dV[rows[sl]], Vn, rtol=0, atol=1e-9)


def test_pack_plan_respects_budgets():
    models = [W.uneven_tree_cell(i, cmin=200, cmax=2000) for i in range(12)]
    L = StaggerLayout(models, threads=256, rows_per_pack=2600)
    P = L.pack_array
    assert np.all(P[:, 3] <= 2600)
    assert np.all(P[:, 1] <= L.threads * L.cv_per_thread)
    assert sorted(c for cells in L.pack_cells for c in cells) == list(range(12))
    assert P[:, 1].sum() == L.n_cv_total


def test_clamp_windows_follow_float64_times():
    dt = 0.025
    n_on = first_step_at(10.0, dt)
    assert n_on * dt >= 10.0 and (n_on - 1) * dt < 10.0
    assert first_step_at(0.0, dt) == 0
    assert first_step_at(1e9, dt) == 2 ** 31 - 1


def test_probes_and_soma_location():
    m = W.single_tree_cell(0)
    L = StaggerLayout([m], probes="soma")
    # x = 0.5 on a 4-CV soma lands on an internal boundary -> owner CV 2 midpoint
    assert L.cells[0].probes.tolist() == [2]
    assert L.clamp_kind == ["mid"]


@pytest.mark.parametrize("max_chain_rows", [16, 2, None])
def test_packed_transposed_solve_matches_dense(max_chain_rows):
    from oracle.dhs import assemble_dense, numeric_state

    models = [W.uneven_tree_cell(i, cmin=200, cmax=300) for i in range(3)]
    L = StaggerLayout(models, rows_per_pack=1200, cv_per_thread=4, max_chain_rows=max_chain_rows)
    dt = 0.025
    rng = np.random.default_rng(4)
    G = rng.uniform(0, 1, L.n_cv_total)
    b = rng.standard_normal(L.n_cv_total)
    rows = cv_rows(L)
    bd = np.zeros(L.n_rows_total)
    bb = np.zeros(L.n_rows_total)
    bd[rows] = 1 + dt * G
    bb[rows] = b
    y = hines_transposed(L, dt, bd, bb)
    yc = hines_transposed_composite(L, dt, bd, bb)
    for i, m in enumerate(models):
        st = StaggeredOracle(m).st
        sl = L.cell_slice(i)
        d, s, lo, up = numeric_state(st, np.zeros((1, st.dynamic_rows.size)), (-G[sl])[None],
                                     np.zeros((1, st.dynamic_rows.size)), dt)
        M = assemble_dense(st, d[0], lo, up)
        rhs = np.zeros(st.n_point)
        rhs[st.dynamic_rows] = b[sl]
        ref = np.linalg.solve(M.T, rhs)
        np.testing.assert_allclose(y[rows[sl]], ref[st.dynamic_rows], rtol=0, atol=1e-10)
        np.testing.assert_allclose(yc[rows[sl]], ref[st.dynamic_rows], rtol=0, atol=1e-10)


def test_slots_are_a_permutation_of_reference_rows():
    models = [W.uneven_tree_cell(i, cmin=200, cmax=300) for i in range(4)]
    L = StaggerLayout(models, rows_per_pack=2000, cv_per_thread=8)
    for p in range(L.n_packs):
        sr = L.slot_rows[p]
        assert sorted(sr.tolist()) == list(range(L.pack_array[p, 3]))


@pytest.mark.parametrize("n_cv", [1, 3, 4])
def test_interior_attachment_matches_oracle(n_cv):
    from tests.helpers import interior_cell

    m = interior_cell(0, n_cv)
    g = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    cvs, _, tree, sched, st = OD.discretize(m.morph, n_cv, 1.0, 100.0)
    assert np.array_equal(g.row_to_point_id, sched.row_to_point_id)
    assert np.array_equal(g.parent_rows, sched.parent_rows)
    assert np.array_equal(g.dynamic_rows, st.dynamic_rows)
    parent, children = OD.node_parent_children(tree)
    assert np.array_equal(g.peel, OD.compute_peel_levels(parent, children))
    for a, b in [(g.diag_axial, st.diag), (g.lowers, st.lowers), (g.uppers, st.uppers), (g.row_cap_uF, st.row_cap_uF)]:
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=0)
    # the interior attachment points are CV midpoints, which stay dynamic rows
    assert g.row_is_mid.sum() == g.n_cv


def test_interior_attachment_packed_hines_matches_oracle_dhs():
    """The kernel's Hines phase (CPU transcription) on packs that mix
    interior-attachment cells with others equals the oracle's DHS solve."""
    from tests.helpers import interior_cell

    models = [interior_cell(s, n) for s, n in ((0, 4), (1, 3), (2, 1))] + [W.uneven_tree_cell(3, cmin=200, cmax=300)]
    L = StaggerLayout(models, rows_per_pack=600, cv_per_thread=2)
    dt = 0.025
    rng = np.random.default_rng(1)
    G = rng.uniform(0, 1, L.n_cv_total)
    V = rng.uniform(-80, 0, L.n_cv_total)
    I = rng.uniform(-5, 5, L.n_cv_total)
    rows, brow = cv_rows(L), bnd_rows(L)
    bd = np.zeros(L.n_rows_total)
    br = np.zeros(L.n_rows_total)
    xv = np.zeros(L.n_rows_total)
    bd[rows] = 1 + dt * G
    br[rows] = dt * I
    xv[rows] = V
    xv[brow] = rng.uniform(-80, 0, len(brow))
    dV = hines_delta(L, dt, bd, br, xv)
    for i, m in enumerate(models):
        st = StaggeredOracle(m).st
        sl = L.cell_slice(i)
        Vn = dhs_solve(st, V[sl][None], (-G[sl])[None], (I[sl] + G[sl] * V[sl])[None], dt)[0]
        np.testing.assert_allclose(V[sl] + dV[rows[sl]], Vn, rtol=0, atol=1e-9)


def test_precomputed_lds_tables_equal_the_kernels_derivation():
    """chain_rec / chain_par / kid_pad / slot_link / slot_chain (copied into
    LDS by the kernel) equal what a workgroup would derive from chain_off,
    chain_top, kid_off and kids, pack by pack."""
    from braincell_amd import workloads as W
    from braincell_amd.runtime import StaggerLayout

    models = [W.uneven_tree_cell(i, cmin=200, cmax=900) for i in range(9)] + [W.single_tree_cell(0)]
    L = StaggerLayout(models, rows_per_pack=2600)
    assert L.n_packs > 1
    rec = L.chain_rec.reshape(-1, 4)
    for P in L.pack_array:
        _, _, row0, n_rows, ch0, n_ch = P[:6]
        kbase = L.kid_off[ch0]
        par = np.full(n_ch, -1)
        link = np.zeros(n_rows, dtype=np.int64)
        chain = np.zeros(n_rows, dtype=np.int64)
        for i in range(n_ch):
            g = ch0 + i
            j0, j1 = L.chain_off[g] - row0, L.chain_off[g + 1] - row0
            q0, nk = L.kid_off[g] - kbase, L.kid_off[g + 1] - L.kid_off[g]
            k0 = L.kids[kbase + q0] if nk > 0 else 0
            k1 = L.kids[kbase + q0 + 1] if nk > 1 else 0
            tp = L.chain_top[g]
            assert tuple(rec[g]) == (j0 | (j1 << 16), (tp & 0xFFFF) | (nk << 16), k0 | (k1 << 16), q0)
            for j in range(j0, j1):
                p = j + 1 if j + 1 < j1 else (tp if tp >= 0 else j)
                link[j] = p | (0x8000 if j == j0 else 0)
                chain[j] = i
            for q in range(nk):
                par[L.kids[kbase + q0 + q]] = i
        nkid = L.kid_off[ch0 + n_ch] - kbase
        assert np.array_equal(L.chain_par[ch0:ch0 + n_ch], par)
        assert np.array_equal(L.kid_pad[ch0:ch0 + nkid], L.kids[kbase:kbase + nkid])
        assert not L.kid_pad[ch0 + nkid:ch0 + n_ch].any()
        assert np.array_equal(L.slot_link[row0:row0 + n_rows].view(np.uint16), link)
        assert np.array_equal(L.slot_chain[row0:row0 + n_rows], chain)


# The chain definition written as a walk over the rows (each bottom row climbs
# while its parent has exactly one child), to check the vectorised builder.
def walk_chains(parent_rows: np.ndarray, max_len: int = None) :
    """Split a row forest (parents given per row; -1 for roots) into chains:
    maximal runs in which every row is the only child of the next, cut from
    the bottom into pieces of at most ``max_len`` rows (a piece's top then
    hangs from the bottom row of the piece above)."""
    parent_rows = np.asarray(parent_rows, dtype=np.int64)
    P = parent_rows.shape[0]
    has_p = parent_rows >= 0
    nchild = np.bincount(parent_rows[has_p], minlength=P)
    bottoms = np.nonzero(nchild != 1)[0]
    chain_of_row = np.full(P, -1, dtype=np.int64)
    rows_list = []
    tops_parent = []
    cap = int(max_len) if max_len else P + 1
    if cap < 1:
        raise ValueError("max_len must be >= 1")
    for b in bottoms.tolist():
        seq = [b]
        r = b
        while True:
            p = parent_rows[r]
            if p < 0 or nchild[p] != 1:
                break
            seq.append(int(p))
            r = p
        top = int(parent_rows[r])
        for q in range(0, len(seq), cap):
            piece = seq[q:q + cap]
            chain_of_row[piece] = len(rows_list)
            rows_list.append(piece)
            tops_parent.append(seq[q + cap] if q + cap < len(seq) else top)
    bottoms = np.asarray([rl[0] for rl in rows_list], dtype=np.int64)
    nch = len(rows_list)
    top_parent = np.asarray(tops_parent, dtype=np.int64)
    parent_chain = np.where(top_parent >= 0, chain_of_row[np.maximum(top_parent, 0)], -1)
    # level: leaves 0; parent = 1 + max(kids).  Process chains from deepest rows
    # upward: a chain's bottom row has a larger peel than every row below it, so
    # ordering by the bottom row index descending (rows are peel-descending in
    # the reference order) visits kids first.
    level = np.zeros(nch, dtype=np.int64)
    ordk = np.argsort(-bottoms, kind="stable")
    for k in ordk.tolist():
        pc = parent_chain[k]
        if pc >= 0:
            level[pc] = max(level[pc], level[k] + 1)
    # sort chains by (level, bottom row)
    perm = np.lexsort((bottoms, level))
    inv = np.empty(nch, dtype=np.int64)
    inv[perm] = np.arange(nch)
    lens = np.asarray([len(rows_list[k]) for k in perm.tolist()], dtype=np.int64)
    off = np.concatenate(([0], np.cumsum(lens)))
    rows = np.concatenate([np.asarray(rows_list[k], dtype=np.int64) for k in perm.tolist()]) if nch else np.zeros(0, np.int64)
    tp = top_parent[perm]
    pch = parent_chain[perm]
    pch_new = np.where(pch >= 0, inv[np.maximum(pch, 0)], -1)
    # kids CSR (kids listed in increasing new chain index)
    kid_cnt = np.bincount(pch_new[pch_new >= 0], minlength=nch)
    kid_off = np.concatenate(([0], np.cumsum(kid_cnt)))
    kids = np.argsort(np.where(pch_new >= 0, pch_new, nch), kind="stable")[: int(kid_cnt.sum())]
    lv = level[perm]
    nlev = int(lv.max()) + 1 if nch else 0
    level_off = np.searchsorted(lv, np.arange(nlev + 1), side="left")
    return dict(off=off, rows=rows, top_parent=tp, kid_off=kid_off, kids=kids,
                  level_off=level_off.astype(np.int64), level=lv)


@pytest.mark.parametrize("max_len", [None, 16, 3, 1])
def test_batch_chains_match_the_row_walk(max_len):
    from braincell_amd.layout import build_chains_batch
    from tests.helpers import interior_cell

    ms = [W.uneven_tree_cell(i, cmin=200, cmax=900) for i in range(8)] + [W.single_tree_cell(0), interior_cell(0),
                                                                          interior_cell(1, 3)]
    gs = [discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra)) for m in ms]
    got = build_chains_batch([g.parent_rows for g in gs], max_len)
    for g, ch in zip(gs, got):
        ref = walk_chains(g.parent_rows, max_len)
        for k, v in ref.items():
            assert np.array_equal(getattr(ch, k), v), k
