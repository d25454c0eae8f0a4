"""SWC reader: the reference reader tests' vectors (``io/swc/reader_test.py``)
restated on branch lengths, radii and attachment coordinates, and the
reference benchmark morphology's published figures (``examples/profiling/
simulator_compare/common.py:30-49``: 217 branches, 868 CVs at 4 per branch,
total length 14438.39058205983 um, branch 0 / 34 / 54 lengths)."""

import hashlib
import os
import textwrap

import numpy as np
import pytest

from braincell_amd.swc import SwcReader, read_swc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "swc")
N144 = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "braincell_amd", "data", "n144.swc")


def _read(tmp_path, body, **kw):
    p = tmp_path / "sample.swc"
    p.write_text(textwrap.dedent(body).strip() + "\n")
    return read_swc(p, **kw)


def test_n144_matches_the_reference_benchmark_figures():
    path = N144
    with open(path, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == "68af99cf829e22005b2177b3d2ff4316bdaec2fcf0f1e4b85f81cdd2a8c01069"
    m = read_swc(path, mode="neuron")
    assert m.n_branches == 217
    assert 4 * m.n_branches == 868
    assert abs(sum(b.total_length for b in m.branches) - 14438.39058205983) <= 1e-6
    for b, L in zip((0, 34, 54), (22.18148613, 38.620829739, 5.313313249)):
        assert abs(m.branches[b].total_length - L) <= 1e-6
    assert m.branches[0].type == "soma"


def test_compressed_tree_and_known_types(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 3 0 10 0 2 1
        3 3 0 20 0 1 2
        4 2 10 0 0 1 1
        5 2 20 0 0 0.5 4
    """)
    assert m.n_branches == 3
    assert m.branch_types() == ["soma", "basal_dendrite", "axon"]
    assert m.branches[0].total_length == pytest.approx(20.0)
    assert m.parent_x[1:] == [0.5, 0.5] and m.child_x[1:] == [0.0, 0.0]
    # neuron mode: a multi-sample child at the soma's midpoint keeps its own first sample
    np.testing.assert_allclose(m.branches[1].lengths, [10.0])
    np.testing.assert_allclose(m.branches[1].radii_proximal, [2.0])


def test_neuromorpho_mode_copies_the_soma_centre(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 3 0 10 0 2 1
        3 3 0 20 0 1 2
        4 2 10 0 0 1 1
        5 2 20 0 0 0.5 4
    """, mode="neuromorpho")
    np.testing.assert_allclose(m.branches[1].lengths, [10.0, 10.0])
    np.testing.assert_allclose(np.r_[m.branches[1].radii_proximal, m.branches[1].radii_distal[-1]], [2.0, 2.0, 1.0])
    np.testing.assert_allclose(m.branches[2].lengths, [10.0, 10.0])


def test_generic_root_that_branches_at_once_extends_itself(tmp_path):
    m = _read(tmp_path, """
        1 2 0 0 0 5 -1
        2 2 10 0 0 4 1
        3 2 -10 0 0 4 1
        4 3 20 0 0 2 2
    """)
    np.testing.assert_allclose(m.branches[0].lengths, [10.0])
    axon_kids = [b for b in range(1, m.n_branches) if m.branches[b].type == "axon"]
    assert len(axon_kids) == 1
    np.testing.assert_allclose(m.branches[axon_kids[0]].lengths, [10.0])
    assert m.parent_x[axon_kids[0]] == 0.0


@pytest.mark.parametrize("side_type", [3, 4])
def test_first_branch_under_the_soma_applies_con2prox(tmp_path, side_type):
    m = _read(tmp_path, f"""
        1 1 0 0 0 10 -1
        2 3 10 0 0 2 1
        3 3 20 0 0 1 2
        4 {side_type} 10 10 0 1 2
    """)
    assert m.parent[1] == 0 and m.parent_x[1] == 0.5
    np.testing.assert_allclose(m.branches[1].lengths, [10.0])
    assert m.parent[2] == 1 and m.parent_x[2] == 0.0
    np.testing.assert_allclose(m.branches[2].lengths, [10.0])
    np.testing.assert_allclose([m.branches[2].radii_proximal[0], m.branches[2].radii_distal[0]], [2.0, 1.0])


def test_single_point_soma(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 3 0 10 0 2 1
    """)
    s = m.branches[0]
    np.testing.assert_allclose(s.lengths, [10.0, 10.0])
    np.testing.assert_allclose(np.r_[s.radii_proximal, s.radii_distal], 10.0)
    d = m.branches[1]
    np.testing.assert_allclose(d.lengths, [10.0])  # from the centre
    assert d.radii_proximal[0] == d.radii_distal[0] == 2.0
    assert m.parent_x[1] == 0.5


def test_special_three_point_soma(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 1 -10 0 0 10 1
        3 1 0 10 0 10 1
        4 3 0 10 0 2 1
    """)
    np.testing.assert_allclose(m.branches[0].lengths, [10.0, 10.0])
    d = m.branches[1]
    np.testing.assert_allclose(d.lengths, [10.0])
    assert d.radii_proximal[0] == d.radii_distal[0] == 2.0
    assert m.parent_x[1] == 0.5


def test_special_three_point_soma_immediate_bifurcation(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 1 -10 0 0 10 1
        3 1 0 10 0 10 1
        4 3 0 20 0 2 1
        5 3 -5 40 10 1 4
        6 3 10 40 10 1 4
    """)
    assert m.parent_x[1] == 0.5 and m.parent[2] == 1 and m.parent_x[2] == 0.0
    np.testing.assert_allclose(m.branches[1].lengths, [np.linalg.norm([-5, 20, 10])])
    np.testing.assert_allclose(m.branches[2].lengths, [np.linalg.norm([10, 20, 10])])


def test_three_point_soma_with_a_side_child_is_a_regular_soma(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 10 -1
        2 1 -10 0 0 10 1
        3 1 0 10 0 10 1
        4 3 -10 10 0 2 2
    """)
    assert sum(t == "soma" for t in m.branch_types()) == 2
    assert m.parent_x[m.branch_types().index("basal_dendrite")] == 1.0


def test_unbranched_multi_point_soma(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 5 -1
        2 1 10 0 0 5 1
        3 1 20 0 0 5 2
        4 3 10 10 0 2 2
        5 3 20 10 0 2 3
    """)
    assert m.n_branches == 3
    assert m.parent_x[1:] == [0.5, 1.0]


def test_branched_multi_point_soma(tmp_path):
    m = _read(tmp_path, """
        1 1 0 0 0 5 -1
        2 1 10 0 0 5 1
        3 1 20 0 0 5 2
        4 1 10 10 0 5 2
        5 3 20 10 0 2 4
    """)
    assert sum(t == "soma" for t in m.branch_types()) == 3
    assert m.parent_x[m.branch_types().index("basal_dendrite")] == 1.0
    m = _read(tmp_path, """
        1 1 0 0 0 5 -1
        2 1 10 0 0 5 1
        3 1 20 0 0 5 2
        4 1 30 0 0 5 3
        5 1 10 10 0 5 2
        6 3 -10 0 0 2 1
        7 3 10 20 0 2 5
        8 3 40 0 0 2 4
    """)
    dends = [b for b in range(m.n_branches) if m.branches[b].type == "basal_dendrite"]
    assert [m.parent_x[b] for b in dends] == [0.0, 1.0, 1.0]
    m = _read(tmp_path, """
        1 1 0 0 100 10 -1
        2 1 0 0 0 10 1
        3 1 10 0 0 10 2
        4 1 20 0 0 10 3
        5 1 -5 0 0 10 2
        6 1 -10 0 0 10 5
        7 3 100 0 0 1 2
    """)
    somas = [b for b in range(m.n_branches) if m.branches[b].type == "soma"]
    assert len(somas) == 3
    np.testing.assert_allclose(m.branches[somas[0]].lengths, [100.0])
    np.testing.assert_allclose(m.branches[somas[1]].lengths, [10.0, 10.0])  # starts at the shared sample
    np.testing.assert_allclose(m.branches[somas[2]].lengths, [5.0, 5.0])


@pytest.mark.parametrize("name", ["single_point_soma", "three_points_soma", "unbranched_soma", "branched_soma",
                                  "branched_soma_1", "branched_dend", "example_tree", "contour_soma"])
def test_bundled_files_read_and_discretise(name):
    from braincell_amd.layout import discretize_cell
    from oracle import discretize as OD

    m = read_swc(os.path.join(HERE, name + ".swc"))
    assert m.n_branches >= 1
    g = discretize_cell(m, np.full(m.n_branches, 3), np.ones(m.n_branches), np.full(m.n_branches, 100.0))
    cvs, _, _, sched, st = OD.discretize(m, 3, 1.0, 100.0)
    assert np.array_equal(g.parent_rows, sched.parent_rows)
    np.testing.assert_allclose(g.cv_area_um2, [c.area_um2 for c in cvs], rtol=1e-12)
    np.testing.assert_allclose(g.diag_axial, st.diag, rtol=1e-12)


def test_branched_soma_file_matches_the_reference_test():
    m = read_swc(os.path.join(HERE, "branched_soma_1.swc"))
    assert m.n_branches == 3
    soma_kids = [b for b in range(1, m.n_branches) if m.parent[b] == 0 and m.branches[b].type == "soma"]
    assert len(soma_kids) == 1 and m.parent_x[soma_kids[0]] == 0.0
    assert m.parent_x[m.branch_types().index("basal_dendrite")] == 1.0


def test_normalisation_warnings_and_errors(tmp_path):
    m, rep = _read(tmp_path, """
        # comment
        10 1 0 0 0 5 -1
        12 7 10 0 0 -1 10
        11 3 5 0 0 1 10
        13 3 5 0 0 1 11
    """, return_report=True)
    codes = rep.codes()
    assert "semantics.unknown_type" in codes and "geometry.radius_positive_double" in codes
    assert "geometry.duplicate_xyzr_node" in codes  # 13 equals 11
    assert "identity.sequential_index" in codes
    assert m.n_branches == 3
    with pytest.raises(ValueError, match="duplicate"):
        _read(tmp_path, """
            1 1 0 0 0 5 -1
            1 3 10 0 0 1 1
        """)
    with pytest.raises(ValueError, match="root"):
        _read(tmp_path, """
            1 1 0 0 0 5 -1
            2 3 10 0 0 1 -1
        """)
    with pytest.raises(ValueError, match="column"):
        _read(tmp_path, """
            1 1 0 0 0 5
        """)
    with pytest.raises(ValueError):
        SwcReader(mode="other")


def test_n144_discretisation_matches_the_oracle():
    from braincell_amd.layout import discretize_cell
    from oracle import discretize as OD

    m = read_swc(N144)
    g = discretize_cell(m, np.full(m.n_branches, 4), np.ones(m.n_branches), np.full(m.n_branches, 1000.0))
    assert g.n_cv == 868
    cvs, _, _, sched, st = OD.discretize(m, 4, 1.0, 1000.0)
    assert np.array_equal(g.row_to_point_id, sched.row_to_point_id)
    assert np.array_equal(g.parent_rows, sched.parent_rows)
    np.testing.assert_allclose(g.cv_area_um2, [c.area_um2 for c in cvs], rtol=1e-12)
    for a, b in [(g.diag_axial, st.diag), (g.lowers, st.lowers), (g.uppers, st.uppers)]:
        np.testing.assert_allclose(a, b, rtol=1e-12)
