"""Shared helpers: run the HIP path and the oracle on the same inputs."""

import numpy as np


def run_gpu(models, n_steps, dt=0.025, precision=32, probes="soma", chunk=None, **layout_kw):
    import torch
    from braincell_amd.runtime import StaggerLayout, StaggerRuntime

    L = StaggerLayout(models, probes=probes, **layout_kw)
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
