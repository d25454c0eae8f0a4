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


# ------------------------------------------------------------ c3-scale cells
C3_STEPS_SHORT, C3_STEPS_LONG = 2400, 8000  # 60 ms and 200 ms at dt 0.025 ms


def _c3_gradient(precision, steps, *, repeat=False):
    """Two c3 region trees (250 branches, ~1000 CVs, their own soma drive of
    0.1-0.5 nA from 10 ms: repetitive spiking), MSE of the soma traces
    against a target run with lognormal-perturbed per-region densities.
    Returns (runtime, target, base densities, d loss / d log density, spike
    counts, and with ``repeat`` a second full forward + reverse pass)."""
    import torch
    from braincell_amd.adjoint import GRAD_CHANNELS, AdjointRuntime

    models = [W.region_tree_cell(i) for i in range(2)]
    base = np.asarray([W.BASE_DENSITIES[k] for k in GRAD_CHANNELS])
    hidden = base[None, None, :] * np.exp(0.3 * np.random.default_rng(5).standard_normal((2, 3, 4)))
    th0 = np.broadcast_to(base, (2, 3, 4)).copy()
    adj = AdjointRuntime(models, 0.025, precision=precision, checkpoint_every=200)
    adj.set_params(torch.as_tensor(hidden))
    target = adj.forward(steps).double().clone()

    def once():
        adj.set_params(torch.as_tensor(th0))
        tr = adj.forward(steps).double()
        gc = adj.backward(2.0 * (tr - target) / tr.numel())
        return tr, gc.clone(), adj.last_Vbar.clone()

    tr, gc, vb = once()
    sens = adj.theta_grad(gc).double().cpu().numpy() * th0
    v = tr.cpu().numpy()
    spikes = [int(((v[:-1, i] < 0) & (v[1:, i] >= 0)).sum()) for i in range(2)]
    second = once()[1:] if repeat else None
    return adj, target, th0, sens, spikes, (gc, vb), second


def _fd_log(adj, target, th0, steps, idx, eps):
    """Central difference of the FP64 loss in log(density) of entry ``idx``."""
    import torch

    def loss(th):
        adj.set_params(torch.as_tensor(th))
        t = adj.forward(steps).double()
        return float(((t - target) ** 2).mean())

    tp, tm = th0.copy(), th0.copy()
    tp[idx] *= np.exp(eps)
    tm[idx] *= np.exp(-eps)
    return (loss(tp) - loss(tm)) / (2 * eps)


def test_gpu_adjoint_c3_cells_fp32_matches_fp64_over_60ms_of_spiking(cuda):
    """2400 steps through 6 and 11 spikes: the FP32 reverse pass agrees with
    the FP64 one to ~2e-5 of the largest log-density gradient (bound 2e-4),
    and the FP64 one with central differences of the FP64 forward (eps 1e-5
    in log density) to ~3e-7 (bound 1e-5)."""
    a64, tgt, th0, s64, spikes, _, _ = _c3_gradient(64, C3_STEPS_SHORT)
    _, _, _, s32, spikes32, _, _ = _c3_gradient(32, C3_STEPS_SHORT)
    assert min(spikes) >= 5 and spikes32 == spikes, (spikes, spikes32)
    scale = np.abs(s64).max()
    err = np.abs(s32 - s64).max() / scale
    print("spikes", spikes, "fp32 vs fp64 / max", err)
    assert err < 2e-4
    for idx in [(0, 0, 0), (0, 1, 2), (1, 0, 1), (1, 2, 3), (0, 2, 2), (1, 1, 0)]:
        fd = _fd_log(a64, tgt, th0, C3_STEPS_SHORT, idx, 1e-5)
        print(idx, s64[idx], fd)
        assert abs(s64[idx] - fd) <= 1e-5 * scale + 1e-5 * abs(fd), (idx, s64[idx], fd)


def test_gpu_adjoint_is_bitwise_repeatable(cuda):
    """Two full forward + reverse passes over 60 ms of c3 cells in FP32 give
    identical class-table gradients and carried voltage adjoints: the
    per-class sums are reduced in a fixed order."""
    import torch

    _, _, _, _, _, (g1, v1), (g2, v2) = _c3_gradient(32, C3_STEPS_SHORT, repeat=True)
    assert torch.equal(g1, g2)
    assert torch.equal(v1, v2)
    assert float(g1.abs().max()) > 0


def test_gpu_adjoint_c3_cells_over_200ms_through_a_mode_switch(cuda):
    """8000 steps (42 spikes on cell 1).  Around 128 ms cell 1's spike train
    passes a mode switch: the derivative of its spike times in the apical CaL
    log density jumps there from ~0.8 ms to ~120 ms (measured with a 1e-7
    relative change), and the apical Na log-density gradient of the loss
    reaches ~5e4, 40x the others.  The FP64 reverse
    pass is the derivative of that map: central differences converge to it
    as eps -> 0 (eps 1e-7, bound 1e-3; measured 8e-7 and 1.3e-4), while at
    eps 1e-5 they cross the switch and differ by up to 100x.  FP32 rounding
    acts as a 6e-8 perturbation there, so cell 1's FP32 gradients are held
    to 10 % only; cell 0 (no switch) to 1e-4."""
    a64, tgt, th0, s64, spikes, _, _ = _c3_gradient(64, C3_STEPS_LONG)
    _, _, _, s32, _, _, _ = _c3_gradient(32, C3_STEPS_LONG)
    assert spikes[1] >= 40, spikes
    for idx in [(1, 2, 3), (1, 2, 0), (1, 0, 1)]:
        fd = _fd_log(a64, tgt, th0, C3_STEPS_LONG, idx, 1e-7)
        print(idx, s64[idx], fd)
        assert abs(s64[idx] - fd) <= 1e-3 * abs(fd), (idx, s64[idx], fd)
    e0 = np.abs(s32[0] - s64[0]).max() / np.abs(s64[0]).max()
    e1 = np.abs(s32[1] - s64[1]).max() / np.abs(s64[1]).max()
    print("fp32 vs fp64 per cell", e0, e1)
    assert e0 < 1e-4 and e1 < 0.1
