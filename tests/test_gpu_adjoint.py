"""GPU hand-written adjoint vs the FP64 oracle adjoint and central finite
differences of the GPU FP64 forward."""

import numpy as np
import pytest

from braincell_amd import workloads as W
from tests.helpers import soma_cv

pytestmark = pytest.mark.gpu


def _setup(n_cells, precision, n_branches=16, K=100):
    from braincell_amd.adjoint import AdjointRuntime

    models = [W.region_tree_cell(i, n_branches=n_branches) for i in range(n_cells)]
    for m in models:
        m.clamps = [(0, 0.5, 1.0, 1e9, 0.3)]  # drive the cells inside the window
    adj = AdjointRuntime(models, 0.025, precision=precision, checkpoint_every=K)
    return models, adj


def test_gpu_adjoint_fp64_matches_oracle_adjoint(cuda):
    import torch
    from oracle.adjoint import adjoint, forward_states
    from oracle.staggered import StaggeredOracle

    steps = 300
    models, adj = _setup(2, 64)
    rng = np.random.default_rng(1)
    wts = rng.standard_normal((steps, 2))
    theta = torch.as_tensor(adj.params(), dtype=torch.float64)
    adj.set_params(theta)
    adj.forward(steps)
    g = adj.theta_grad(adj.backward(torch.as_tensor(wts, device="cuda"))).cpu().numpy()  # [cells, regions, 4]
    for i, m in enumerate(models):
        o = StaggeredOracle(m)
        o.reset_state(m.V0)
        states = forward_states(o, 0.025, steps)
        o2 = StaggeredOracle(m)
        o2.reset_state(m.V0)
        gr = adjoint(o2, states, 0.025, wts[:, i], soma_cv(m))
        types = [b.type for b in m.morph.branches]
        for r, reg in enumerate(W.REGIONS):
            cvm = np.asarray([t == reg for t in types])[o.branch_of]
            for k, nm in enumerate(("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")):
                ref = gr[nm][cvm].sum() * 1e3  # per S/cm^2
                assert abs(g[i, r, k] - ref) <= 1e-6 * max(1.0, abs(ref)) + 1e-5 * abs(ref), (i, reg, nm, g[i, r, k], ref)


def test_gpu_adjoint_fp64_matches_fd_and_fp32_is_close(cuda):
    import torch

    steps = 400
    models, adj = _setup(2, 64, K=150)
    rng = np.random.default_rng(2)
    target = torch.as_tensor(rng.uniform(-70, -60, (steps, 2)), device="cuda", dtype=torch.float64)
    from braincell_amd.adjoint import StaggerTraceFunction

    theta = torch.as_tensor(adj.params(), dtype=torch.float64, device="cuda").requires_grad_(True)
    tr = StaggerTraceFunction.apply(theta, adj, steps)
    loss = ((tr - target) ** 2).mean()
    loss.backward()
    g = theta.grad.detach().cpu().numpy()

    def L(th):
        with torch.no_grad():
            adj.set_params(th)
            t = adj.forward(steps)
            return float(((t - target) ** 2).mean())

    th0 = theta.detach().clone()
    checked = 0
    for (i, r, k) in [(0, 0, 0), (0, 0, 2), (1, 1, 1), (1, 2, 2), (0, 2, 3)]:
        if adj.cls_rows[i, r] < 0:
            continue
        h = 1e-4 * float(th0[i, r, k])
        tp, tm = th0.clone(), th0.clone()
        tp[i, r, k] += h
        tm[i, r, k] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        scale = np.abs(g).max()
        print((i, r, k), g[i, r, k], fd)
        assert abs(g[i, r, k] - fd) <= 1e-4 * scale + 1e-3 * abs(fd)
        checked += 1
    assert checked >= 3
    # FP32 adjoint close to FP64
    _, adj32 = _setup(2, 32, K=150)
    th32 = th0.to(torch.float32).requires_grad_(True)
    tr32 = StaggerTraceFunction.apply(th32, adj32, steps)
    ((tr32 - target.float()) ** 2).mean().backward()
    g32 = th32.grad.double().cpu().numpy()
    rel = np.abs(g32 - g).max() / np.abs(g).max()
    print("fp32 vs fp64 gradient, max err / max |g|:", rel)
    assert rel < 1e-2


@pytest.mark.parametrize("threads,cvpt", [(256, 4), (128, 8)])
def test_gpu_adjoint_fp64_large_trees_match_oracle_adjoint(cuda, threads, cvpt):
    """Two ~1000-CV trees (chain trees of depth ~20: several pointer-jumping
    rounds in the reverse solve), at 4 and 8 CVs per thread in FP64 (the
    reverse kernel then reads the recorded states where it uses them instead
    of carrying them in registers), over 240 steps in checkpoint segments of
    100, 100 and 40."""
    import torch
    from braincell_amd.adjoint import AdjointRuntime
    from oracle.adjoint import adjoint, forward_states
    from oracle.staggered import StaggeredOracle

    steps = 240
    models = [W.region_tree_cell(i, n_branches=250) for i in range(2)]
    for m in models:
        m.clamps = [(0, 0.5, 1.0, 1e9, 0.3)]
    adj = AdjointRuntime(models, 0.025, precision=64, checkpoint_every=100, threads=threads)
    assert adj.L.cv_per_thread == cvpt and adj.L.max_levels > 10
    wts = np.random.default_rng(3).standard_normal((steps, 2))
    adj.set_params(torch.as_tensor(adj.params(), dtype=torch.float64))
    adj.forward(steps)
    g = adj.theta_grad(adj.backward(torch.as_tensor(wts, device="cuda"))).cpu().numpy()
    for i, m in enumerate(models):
        o = StaggeredOracle(m)
        o.reset_state(m.V0)
        states = forward_states(o, 0.025, steps)
        o2 = StaggeredOracle(m)
        o2.reset_state(m.V0)
        gr = adjoint(o2, states, 0.025, wts[:, i], soma_cv(m))
        types = [b.type for b in m.morph.branches]
        for r, reg in enumerate(W.REGIONS):
            cvm = np.asarray([t == reg for t in types])[o.branch_of]
            for k, nm in enumerate(("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")):
                ref = gr[nm][cvm].sum() * 1e3  # per S/cm^2
                assert abs(g[i, r, k] - ref) <= 1e-6 * max(1.0, abs(ref)) + 1e-5 * abs(ref), (i, reg, nm, g[i, r, k], ref)
