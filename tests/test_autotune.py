"""Pack-plan autotuner: candidate plans, repacking, and (GPU) equal results
under every plan."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from braincell_amd.autotune import candidate_plans
from braincell_amd.runtime import StaggerLayout


def _batch(n=24):
    return [W.uneven_tree_cell(i, cmin=200, cmax=900) for i in range(n)]


def test_repacked_layout_equals_a_fresh_one():
    models = _batch()
    base = StaggerLayout(models, threads=256)
    for t, c, r in [(128, 8, 1500), (256, 4, 1100)]:
        a = base.repacked(threads=t, cv_per_thread=c, rows_per_pack=r)
        b = StaggerLayout(models, threads=t, cv_per_thread=c, rows_per_pack=r)
        assert a.n_packs == b.n_packs and a.pack_cells == b.pack_cells
        for k in ("cv_info", "bnd_info", "chain_off", "chain_top", "kid_off", "kids", "level_off", "slot_lo"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert base.n_packs == StaggerLayout(models, threads=256).n_packs  # the base is untouched


def test_candidate_plans_fit_every_cell_and_lds_grows_with_rows():
    models = _batch()
    base = StaggerLayout(models)
    ncv = max(c.geom.n_cv for c in base.cells)
    plans = candidate_plans(base)
    assert len(plans) >= 4 and len(set(plans)) == len(plans)
    by_rows = {}
    for t, c, r in plans:
        assert t * c >= ncv
        L = base.repacked(threads=t, cv_per_thread=c, rows_per_pack=r)
        assert all(sum(L.cells[i].geom.n_point for i in p) <= r for p in L.pack_cells)
        by_rows.setdefault(t, []).append((r, L.lds_bytes()))
    for t, v in by_rows.items():
        v.sort()
        assert all(b[1] >= a[1] for a, b in zip(v, v[1:]))


@pytest.mark.gpu
def test_every_plan_gives_the_same_trajectories(cuda, tmp_path, monkeypatch):
    import torch
    from braincell_amd.autotune import autotune
    from braincell_amd.runtime import StaggerRuntime

    monkeypatch.setenv("BRAINCELL_AMD_CACHE", str(tmp_path))
    models = _batch(40)
    base = StaggerLayout(models)
    best, table = autotune(base, 0.025, steps=20, reps=1)
    assert len(table) >= 3

    def soma(L):
        rt = StaggerRuntime(L, 0.025, precision=32)
        tr = rt.run(300, chunk=100)
        torch.cuda.synchronize()
        cols = [int(L.cell_probe0[i]) for i in range(len(models))]
        return tr.cpu().numpy()[:, cols]

    ref = soma(base)
    assert any(len(p) == 4 for p, _ in table)  # chain cut lengths were searched
    assert {p[4] for p, _ in table if len(p) == 5} == {"sweep"}  # and the other back substitution
    for plan, _ in table:
        t, c, r = plan[:3]
        if len(plan) == 5:
            # the sequential back substitution sums the chain tree in another order
            L = base.rechained(plan[3], threads=t, cv_per_thread=c, rows_per_pack=r).with_backsub(plan[4])
            assert L.step_flags == 8
            got = soma(L)
            assert np.max(np.abs(got - ref)) < 1e-3, plan
        elif len(plan) == 4:
            # another chain cut groups the elimination's rounding differently
            got = soma(base.rechained(plan[3], threads=t, cv_per_thread=c, rows_per_pack=r))
            assert np.max(np.abs(got - ref)) < 1e-3, plan
        else:
            got = soma(base.repacked(threads=t, cv_per_thread=c, rows_per_pack=r))
            # per-cell arithmetic does not depend on the plan (kernels compiled for
            # other register-slot counts may contract differently: 1e-4 mV)
            assert np.max(np.abs(got - ref)) < 1e-4, plan
    # the cached plan comes back without timing again
    again, _ = autotune(base, 0.025, steps=20, reps=1)
    assert (again.threads, again.cv_per_thread, again.rows_per_pack, again.max_chain_rows, again.backsub) == (
        best.threads, best.cv_per_thread, best.rows_per_pack, best.max_chain_rows, best.backsub)


def test_backsub_variant_sets_the_step_flag():
    base = StaggerLayout(_batch(4), threads=256)
    assert base.backsub == "jump" and base.step_flags == 0
    sw = base.with_backsub("sweep")
    assert sw.step_flags == 8 and base.step_flags == 0  # the base is untouched
    assert sw.repacked(rows_per_pack=3000).backsub == "sweep"
    assert sw.rechained(12).backsub == "sweep"
    assert sw.summary()["backsub"] == "sweep"
    with pytest.raises(ValueError):
        base.with_backsub("doubling")


def test_rechained_layout_equals_a_fresh_one():
    models = _batch(12)
    base = StaggerLayout(models, threads=256)
    for S in (12, 4, None):
        a = base.rechained(S, rows_per_pack=1500)
        b = StaggerLayout(models, threads=256, rows_per_pack=1500, max_chain_rows=S)
        assert a.max_chain_rows == b.max_chain_rows and a.pack_cells == b.pack_cells
        for k in ("cv_info", "bnd_info", "chain_off", "chain_top", "kid_off", "kids", "level_off", "chain_rec",
                  "chain_work", "slot_link"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), (S, k)
    assert base.max_chain_rows == 16 and base.cells[0].chains is not a.cells[0].chains
