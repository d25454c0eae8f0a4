"""The FP64 oracle adjoint (reverse of the staggered step) against central
finite differences of the oracle's loss (CPU)."""

import copy

import numpy as np

from braincell_amd import workloads as W
from oracle.adjoint import adjoint, forward_states
from oracle.staggered import StaggeredOracle


def test_oracle_adjoint_matches_central_fd():
    dt, steps = 0.025, 300
    m = W.region_tree_cell(0, n_branches=12)
    m.clamps = [(0, 0.5, 1.0, 1e9, 0.3)]  # drive the cell inside the window
    rng = np.random.default_rng(0)
    wts = rng.standard_normal(steps)

    def loss(model):
        o = StaggeredOracle(model)
        o.reset_state(model.V0)
        return float(np.dot(wts, o.run(dt, steps, probe_cvs=[2])[:, 0]))

    o = StaggeredOracle(m)
    o.reset_state(m.V0)
    states = forward_states(o, dt, steps)
    o2 = StaggeredOracle(m)
    o2.reset_state(m.V0)
    grad = adjoint(o2, states, dt, wts, 2)
    types = [b.type for b in m.morph.branches]
    adj_all, fd_all = [], []
    for name in ["Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008"]:
        for reg in W.REGIONS:
            bmask = np.asarray([t == reg for t in types])
            if not bmask.any():
                continue
            g0 = m.channels[name]["g_branch"][bmask][0] * 1e3
            h = 1e-4 * g0
            mp, mm = copy.deepcopy(m), copy.deepcopy(m)
            mp.channels[name]["g_branch"] = np.where(bmask, (g0 + h) * 1e-3, m.channels[name]["g_branch"])
            mm.channels[name]["g_branch"] = np.where(bmask, (g0 - h) * 1e-3, m.channels[name]["g_branch"])
            fd_all.append((loss(mp) - loss(mm)) / (2 * h))
            adj_all.append(grad[name][bmask[o.branch_of]].sum())
    adj_all, fd_all = np.asarray(adj_all), np.asarray(fd_all)
    scale = np.abs(fd_all).max()
    err = np.abs(adj_all - fd_all)
    assert np.all(err <= 1e-4 * scale + 1e-3 * np.abs(fd_all)), (adj_all, fd_all)
