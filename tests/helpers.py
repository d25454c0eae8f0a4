"""Shared helpers: run the HIP path and the oracle on the same inputs."""

import numpy as np


def run_gpu(models, n_steps, dt=0.025, precision=32, probes="soma", chunk=None, backsub=None, **layout_kw):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    L = StaggerLayout(models, probes=probes, **layout_kw)
    if backsub is not None:
        L = L.with_backsub(backsub)
    rt = StaggerRuntime(L, dt, precision=precision)
    tr = rt.run(n_steps, chunk=chunk)
    torch.cuda.synchronize()
    rt.check_status()
    return tr.double().cpu().numpy(), rt


def run_oracle(model, n_steps, dt=0.025, probe_cvs=None):
    from oracle.staggered import StaggeredOracle

    o = StaggeredOracle(model)
    o.reset_state(model.V0)
    return o.run(dt, n_steps, probe_cvs=probe_cvs), o


def soma_cv(model):
    from braincell_amd.layout import discretize_cell

    g = discretize_cell(model.morph, model.n_cv_branch(), model.per_branch(model.cm), model.per_branch(model.ra))
    kind, idx = g.locate(0, 0.5, model.n_cv_branch())
    return idx


def spike_times(v, dt, threshold=0.0):
    """Upward crossings of ``threshold`` (get_spike, braincell/_base.py:421-431):
    step n spikes when V_{n-1} < th <= V_n; time = (n + 1) dt."""
    v = np.asarray(v)
    idx = np.nonzero((v[:-1] < threshold) & (v[1:] >= threshold))[0] + 1
    return (idx + 1) * dt


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


def interior_cell(seed=0, n_cv=4):
    """A cell whose children attach inside their parents (0 < parent_x < 1),
    at branch ends, at x = 0 of a parent and with a reversed walk
    (child_x = 1); Na/KDR/leak everywhere, CaL and a Ca pool on dendrites,
    a soma clamp from 1 ms."""
    from braincell_amd.model import CellModel
    from braincell_amd.morph import Branch, Morphology
    from braincell_amd.workloads import _ca_pool

    rng = np.random.default_rng(seed)
    soma = Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma")
    m = Morphology.from_root(soma)

    def dend():
        return Branch.from_lengths(lengths=[float(rng.uniform(30, 150))],
                                   radii=[float(rng.uniform(0.6, 1.0)), float(rng.uniform(0.3, 0.6))], type="dendrite")

    m.attach(0, dend(), parent_x=0.5)  # an internal CV boundary of the soma -> owner CV
    m.attach(0, dend(), parent_x=0.3)
    m.attach(0, dend(), parent_x=1.0)
    m.attach(1, dend(), parent_x=0.37)
    m.attach(1, dend(), parent_x=1.0)
    m.attach(2, dend(), parent_x=0.0)
    m.attach(3, dend(), parent_x=0.8, child_x=1.0)
    m.attach(4, dend(), parent_x=0.61)
    m.attach(7, dend(), parent_x=1.0)
    nb = m.n_branches
    dmask = np.asarray([b.type != "soma" for b in m.branches])
    return CellModel(
        morph=m, cv_per_branch=n_cv, cm=1.0, ra=100.0, V0=-70.0,
        channels={
            "Na_Ba2002": {"g_branch": np.full(nb, 0.12), "params": {}},
            "KDR_Ba2002": {"g_branch": np.full(nb, 0.036), "params": {}},
            "IL": {"g_branch": np.full(nb, 0.0003), "params": {"E": -70.0}},
            "CaL_IS2008": {"g_branch": np.where(dmask, 0.002, 0.0), "params": {}},
        },
        ca=_ca_pool(dmask),
        clamps=[(0, 0.5, 1.0, 1.0e9, 0.6)],
    ).validate()
