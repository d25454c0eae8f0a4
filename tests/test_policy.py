"""CV policies: the reference's policy tests (``_discretization/
policy_test.py``) restated, plus the d_lambda rule against its closed form
and the lowered layouts against the oracle."""

import numpy as np
import pytest

from braincell_amd.morph import Branch, Morphology
from braincell_amd.policy import (CompositeByTypePolicy, CVPerBranch, CVPerBranchList, CVPolicy, CVPolicyByTypeRule,
                                  DLambda, MaxCVLen)


def _two_branch():
    m = Morphology.from_root(Branch.from_lengths(lengths=[100.0], radii=[10.0, 8.0], type="soma"))
    m.attach(0, Branch.from_lengths(lengths=[45.0], radii=[2.0, 1.0], type="basal_dendrite"))
    return m


def _three_branch():
    m = Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[8.0, 8.0], type="soma"))
    m.attach(0, Branch.from_lengths(lengths=[30.0], radii=[2.0, 1.5], type="basal_dendrite"))
    m.attach(0, Branch.from_lengths(lengths=[40.0], radii=[2.5, 1.0], type="apical_dendrite"))
    return m


def _mixed():
    m = Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[8.0, 8.0], type="soma"))
    m.attach(0, Branch.from_lengths(lengths=[80.0], radii=[1.0, 0.8], type="axon"))
    m.attach(0, Branch.from_lengths(lengths=[30.0], radii=[2.0, 1.5], type="basal_dendrite"))
    m.attach(0, Branch.from_lengths(lengths=[45.0], radii=[2.5, 1.2], type="apical_dendrite"))
    return m


def test_base_policy_is_abstract():
    with pytest.raises(NotImplementedError):
        CVPolicy().counts(_two_branch())


def test_cv_per_branch_and_list():
    t = _three_branch()
    assert CVPerBranch(3).resolve_cv_bounds(t) == tuple(((0.0, 1 / 3), (1 / 3, 2 / 3), (2 / 3, 1.0)) for _ in range(3))
    assert CVPerBranchList((1, 3, 5)).counts(t).tolist() == [1, 3, 5]
    with pytest.raises(ValueError, match="branch count"):
        CVPerBranchList((1, 2)).counts(t)
    with pytest.raises(TypeError):
        CVPerBranchList((1, 2.0, 3))
    with pytest.raises(ValueError):
        CVPerBranchList((1, 0, 3))
    with pytest.raises(TypeError):
        CVPerBranch(True)
    with pytest.raises(ValueError):
        CVPerBranch(0)


def test_max_cv_len():
    t = _two_branch()
    p = MaxCVLen(20.0)
    assert p.resolve_cv_bounds(t) == (((0.0, 0.2), (0.2, 0.4), (0.4, 0.6), (0.6, 0.8), (0.8, 1.0)),
                                      ((0.0, 1 / 3), (1 / 3, 2 / 3), (2 / 3, 1.0)))
    soma = Morphology.from_root(Branch.from_lengths(lengths=[10.0], radii=[3.0, 2.0], type="soma"))
    assert MaxCVLen(5.0).counts(soma).tolist() == [3]  # 2 promoted to odd
    assert MaxCVLen(5.0, keep_odd=False).counts(soma).tolist() == [2]
    assert MaxCVLen(10.0 + 5e-7).counts(soma).tolist() == [1]
    # every CV stays below the maximum
    for L in (7.0, 33.3, 100.0, 250.0):
        m = Morphology.from_root(Branch.from_lengths(lengths=[L], radii=[1.0, 1.0], type="dendrite"))
        n = int(MaxCVLen(12.5, keep_odd=False).counts(m)[0])
        assert L / n <= 12.5 + 1e-9 and (n == 1 or L / (n - 1) > 12.5)
    with pytest.raises(ValueError):
        MaxCVLen(0.0)
    with pytest.raises(TypeError):
        MaxCVLen("20")


def test_d_lambda_closed_form():
    soma = Morphology.from_root(Branch.from_lengths(lengths=[78.0], radii=[1.0, 1.0], type="soma"))
    assert DLambda(0.1, keep_odd=False).counts(soma).tolist() == [2]
    assert DLambda(0.1).counts(soma).tolist() == [3]
    # lambda_f = 1e5 sqrt(d / (4 pi f Ra cm)) um; a multi-segment branch sums L / lambda_f
    br = Branch.from_lengths(lengths=[50.0, 120.0, 30.0], radii_proximal=[2.0, 1.5, 1.0],
                             radii_distal=[1.5, 1.0, 0.5], type="dendrite")
    m = Morphology.from_root(br)
    f, ra, cm = 250.0, 150.0, 0.8
    lam = 1e5 * np.sqrt((br.radii_proximal + br.radii_distal) / (4 * np.pi * f * ra * cm))
    n = int(np.ceil(np.sum(br.lengths / lam) / 0.05 - 1e-9))
    assert DLambda(0.05, frequency=f, keep_odd=False).counts(m, cm_branch=[cm], ra_branch=[ra]).tolist() == [n]
    # a branch with larger Ra is electrotonically longer
    t = _two_branch()
    c = DLambda(0.01, keep_odd=False).counts(t, cm_branch=[1.0, 1.0], ra_branch=[100.0, 400.0])
    c0 = DLambda(0.01, keep_odd=False).counts(t)
    assert c[1] > c0[1] and c[0] == c0[0]
    with pytest.raises(ValueError):
        DLambda(0.0)
    with pytest.raises(ValueError):
        DLambda(0.1, frequency=-1.0)


def test_composite_by_type():
    t = _mixed()
    p = CompositeByTypePolicy(
        rules=(CVPolicyByTypeRule(("soma",), CVPerBranch(1)),
               CVPolicyByTypeRule(("basal_dendrite", "apical_dendrite"), MaxCVLen(20.0, keep_odd=False))),
        default_policy=CVPerBranch(2))
    assert p.counts(t).tolist() == [1, 2, 2, 3]
    last = CompositeByTypePolicy(rules=(CVPolicyByTypeRule(("axon",), CVPerBranch(1)),
                                        CVPolicyByTypeRule(("axon",), CVPerBranch(3))),
                                 default_policy=CVPerBranch(2))
    assert last.counts(t)[1] == 3
    with pytest.raises(TypeError):
        CompositeByTypePolicy(rules=(), default_policy=object())
    with pytest.raises(ValueError):
        CVPolicyByTypeRule((), CVPerBranch(1))
    with pytest.raises(ValueError):
        CVPolicyByTypeRule(("dendrites",), CVPerBranch(1))


def test_cell_lowering_uses_the_policy_and_matches_the_oracle():
    from braincell_amd.cell import Cell
    from braincell_amd.layout import discretize_cell
    from braincell_amd.mech import CableProperty, Channel
    from oracle import discretize as OD

    t = _mixed()
    cell = Cell(t, cv_policy=MaxCVLen(9.0))
    cell.paint("all", CableProperty(axial_resistivity=120.0), Channel("IL", g_max=0.0003))
    m = cell.lower()
    assert m.n_cv_branch().tolist() == MaxCVLen(9.0).counts(t).tolist() == [3, 9, 5, 5]
    g = discretize_cell(m.morph, m.n_cv_branch(), m.per_branch(m.cm), m.per_branch(m.ra))
    cvs, _, _, sched, st = OD.discretize(t, m.n_cv_branch().tolist(), 1.0, 120.0)
    assert np.array_equal(g.parent_rows, sched.parent_rows)
    np.testing.assert_allclose(g.diag_axial, st.diag, rtol=1e-13)
    with pytest.raises(TypeError):
        Cell(t, cv_policy=4)
