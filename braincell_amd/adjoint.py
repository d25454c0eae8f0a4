"""Gradients of probe traces with respect to channel densities.

The reference differentiates its jitted scan with JAX autodiff
(``_multi_compartment/run.py:108-134``).  Here the reverse pass is the
hand-written adjoint kernel (``bc_stagger_adjoint``) with checkpointing:
the forward keeps the state at every ``checkpoint_every``-th step; the
backward replays each segment once with recording switched on and runs the
adjoint kernel over it, last segment first.  ``StaggerTraceFunction`` exposes
this as a ``torch.autograd.Function`` of per-region densities.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._native import AdjointDesc, check, lib
from .runtime import REGION_CODES, StaggerLayout, StaggerRuntime

__all__ = ["AdjointRuntime", "StaggerTraceFunction", "GRAD_CHANNELS"]

GRAD_CHANNELS = ("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")
GRAD_COLS = (1, 2, 3, 5)  # class-table columns of those densities


class AdjointRuntime:
    """Forward with checkpoints and segment-wise adjoint for one batch."""

    def __init__(self, models, dt: float, *, precision: int = 32, checkpoint_every: int = 200,
                 regions=("soma", "basal_dendrite", "apical_dendrite"), **layout_kw):
        import torch

        self.torch = torch
        self.L = StaggerLayout(models, region_classes=True, **layout_kw)
        self.rt = StaggerRuntime(self.L, dt, precision=precision)
        if not lib().bc_adjoint_supported(precision, self.L.chset, self.L.cv_per_thread):
            raise RuntimeError(f"no adjoint kernel for channel set {hex(self.L.chset)}")
        self.K = int(checkpoint_every)
        self.regions = tuple(regions)
        codes = [REGION_CODES.index(r) for r in self.regions]
        self.cls_rows = self.L.cell_region_cls[:, codes]  # [cells, regions]
        self.ns = len(self.L.slots)
        self.ckpt = []
        self.n_steps = 0

    # --------------------------------------------------------------- params
    def set_params(self, theta):
        """theta [cells, regions, 4] densities in S/cm^2 (Na, KDR, IL, CaL)."""
        t = self.torch
        th = theta.detach().to(self.rt.device, self.rt.dtype)
        tab = self.rt.cls_tab.view(-1, 12)
        ok = t.as_tensor(self.cls_rows >= 0, device=self.rt.device)
        rows = t.as_tensor(np.maximum(self.cls_rows, 0), device=self.rt.device)
        for k, col in enumerate(GRAD_COLS):
            vals = th[:, :, k] * 1e3
            cur = tab[rows, col]
            tab[rows, col] = t.where(ok, vals, cur)

    def params(self):
        tab = self.rt.cls_tab.view(-1, 12).double().cpu().numpy()
        out = np.zeros(self.cls_rows.shape + (4,))
        for k, col in enumerate(GRAD_COLS):
            out[:, :, k] = np.where(self.cls_rows >= 0, tab[np.maximum(self.cls_rows, 0), col], 0.0) * 1e-3
        return out

    # -------------------------------------------------------------- forward
    def forward(self, n_steps: int):
        rt = self.rt
        rt.reset_state()
        trace = rt.new_trace(n_steps)
        self.ckpt = []
        done = 0
        while done < n_steps:
            k = min(self.K, n_steps - done)
            self.ckpt.append((done, k, rt.V.clone(), rt.S.clone(), rt.Vbnd.clone()))
            rt.run(k, trace=trace[done:done + k])
            done += k
        self.n_steps = n_steps
        return trace

    # ------------------------------------------------------------- backward
    def backward(self, trace_bar, events=None):
        """Reverse pass for the loss gradient ``trace_bar`` (d loss / d trace);
        returns the class-table gradient.  ``events``, a list, receives a
        (start, end) pair of timing events around every reverse launch."""
        t = self.torch
        rt = self.rt
        ncv = self.L.n_cv_total
        if trace_bar.shape != (self.n_steps, self.L.n_probes):
            raise ValueError("trace_bar must have the trace's shape")
        trace_bar = trace_bar.to(rt.device, rt.dtype).contiguous()
        Vbar = t.zeros(ncv, dtype=rt.dtype, device=rt.device)
        Sbar = t.zeros(max(1, self.ns * ncv), dtype=rt.dtype, device=rt.device)
        gcls = t.zeros(self.L.cls_tab.shape[0] * 12, dtype=rt.dtype, device=rt.device)
        rec = t.empty((self.K + 1) * (self.ns + 1) * ncv, dtype=rt.dtype, device=rt.device)
        scratch = rt.new_trace(self.K)
        V0, S0, B0, idx0 = rt.V.clone(), rt.S.clone(), rt.Vbnd.clone(), rt.step_index
        for (start, k, V, S, Vb) in reversed(self.ckpt):
            rt.V.copy_(V)
            rt.S.copy_(S)
            rt.Vbnd.copy_(Vb)
            rt.step_index = start
            rt.run(k, trace=scratch[:k], rec=rec)
            f = rt.desc(k, scratch, step0=start)
            d = AdjointDesc()
            d.layout = ctypes.cast(ctypes.pointer(f), ctypes.c_void_p)
            d.n_steps = k
            d.rec = rec.data_ptr()
            d.trace_bar = trace_bar[start:start + k].data_ptr()
            d.Vbar, d.Sbar, d.cls_grad = Vbar.data_ptr(), Sbar.data_ptr(), gcls.data_ptr()
            if events is not None:
                ev = (t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True))
                ev[0].record()
            check(lib().bc_stagger_adjoint(ctypes.byref(d), rt.stream_handle()))
            if events is not None:
                ev[1].record()
                events.append(ev)
        rt.V.copy_(V0)
        rt.S.copy_(S0)
        rt.Vbnd.copy_(B0)
        rt.step_index = idx0
        self.last_Vbar, self.last_Sbar = Vbar, Sbar
        return gcls.view(-1, 12)

    def theta_grad(self, gcls):
        """Class-table gradient (per mS/cm^2) -> [cells, regions, 4] per S/cm^2."""
        t = self.torch
        rows = t.as_tensor(np.maximum(self.cls_rows, 0), device=gcls.device)
        ok = t.as_tensor(self.cls_rows >= 0, device=gcls.device)
        cols = [gcls[rows, c] * 1e3 for c in GRAD_COLS]
        g = t.stack(cols, dim=-1)
        return t.where(ok[..., None], g, t.zeros_like(g))


class StaggerTraceFunction:
    """``apply(theta, adj, n_steps)`` -> probe trace, differentiable in theta."""

    @staticmethod
    def apply(theta, adj: AdjointRuntime, n_steps: int):
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, th):
                adj.set_params(th)
                tr = adj.forward(n_steps)
                return tr.to(th.dtype)

            @staticmethod
            def backward(ctx, gtr):
                g = adj.theta_grad(adj.backward(gtr.to(adj.rt.dtype)))
                return g.to(theta.dtype)

        return _F.apply(theta)
