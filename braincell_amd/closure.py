"""Learned gating closure for the staggered step at large dt.

A per-compartment MLP 8 -> 64 -> 64 -> 5 (tanh) corrects the exponential-
Euler update of the five gates (Na p, q; KDR p; CaL p, q) of the
Na/KDR/IL/CaL/CalciumDetailed model: with delta the plain increment of a
gate at V_{n+1} and c the network output, the gate moves by delta (1 + c),
so a cell at rest (delta = 0) cannot drift whatever the network says.  Everything numeric runs in this
package's kernels (``bc_closure.hip``):

* training pairs come from plain fine-dt trajectories of the staggered
  kernel (``bc_closure_data`` computes the targets with the kernel's own gate
  update);
* the network is trained on the GPU by ``bc_closure_grad`` (FP32 forward and
  backward) and ``bc_closure_adam``;
* inference is BF16 MFMA with FP32 accumulation, fused with feature
  extraction and the state correction (``bc_closure_apply``), after each
  coarse plain step.

The plain method judges the closure: ``calibrate`` runs a window of every
cell with the closure at the coarse dt and with plain fine dt, and cells
whose soma RMSE exceeds the threshold fall back to plain dt = 0.025 ms
(``ClosureBatch``).  No counterpart in the reference.
"""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._native import check, lib
from .runtime import StaggerLayout, StaggerRuntime
from .workloads import CLAMP_FOREVER

__all__ = ["NPARAM", "ClosureNet", "ClosureData", "generate_data", "train", "ClosureStepper", "ClosureBatch",
           "with_clamp", "soma_rmse", "spike_times_ms", "spike_error"]

NPARAM = 5061
NIN, NH, NOUT = 8, 64, 5
CHSET = 0x117  # Na | KDR | IL | CaL | CalciumDetailed


class _DataDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ld", ctypes.c_int64), ("dt", ctypes.c_double), ("chan_const", ctypes.c_void_p),
                ("V1", ctypes.c_void_p), ("S0", ctypes.c_void_p), ("S1", ctypes.c_void_p), ("X", ctypes.c_void_p),
                ("D", ctypes.c_void_p), ("Y", ctypes.c_void_p)]


class _TrainDesc(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int64), ("max_blocks", ctypes.c_int32), ("X", ctypes.c_void_p),
                ("Dn", ctypes.c_void_p), ("Yn", ctypes.c_void_p), ("idx", ctypes.c_void_p), ("P", ctypes.c_void_p), ("partial", ctypes.c_void_p),
                ("loss_part", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("loss", ctypes.c_void_p)]


class _AdamDesc(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int64), ("lr", ctypes.c_double), ("beta1", ctypes.c_double),
                ("beta2", ctypes.c_double), ("eps", ctypes.c_double), ("P", ctypes.c_void_p), ("m", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("grad", ctypes.c_void_p)]


class _FwdDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("mfma", ctypes.c_int32), ("X", ctypes.c_void_p), ("P", ctypes.c_void_p),
                ("norm", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class _ApplyDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ld", ctypes.c_int64), ("dt", ctypes.c_double), ("chan_const", ctypes.c_void_p),
                ("V1", ctypes.c_void_p), ("S", ctypes.c_void_p), ("P", ctypes.c_void_p), ("norm", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("flag_tol", ctypes.c_double)]


SKIP_GATES = 1  # BC_STEP_SKIP_GATES


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def with_clamp(model, amp_nA: float, start_ms: float = 2.0):
    """The model with one soma clamp of ``amp_nA`` from ``start_ms`` on."""
    return dataclasses.replace(model, clamps=[(0, 0.5, float(start_ms), CLAMP_FOREVER, float(amp_nA))])


# ----------------------------------------------------------------- network
class ClosureNet:
    """Parameters (FP32, BF16-MFMA at inference) and normalisation."""

    def __init__(self, P, norm, device="cuda"):
        import torch

        self.P = torch.as_tensor(np.asarray(P, dtype=np.float32), device=device).contiguous()
        self.norm = torch.as_tensor(np.asarray(norm, dtype=np.float32), device=device).contiguous()
        if self.P.numel() != NPARAM or self.norm.numel() != 2 * NIN + NOUT:
            raise ValueError("closure parameters must hold 5061 values and norm 21")

    @staticmethod
    def init_params(seed: int = 0, out_gain: float = 0.01) -> np.ndarray:
        """Glorot-scaled random hidden layers, near-zero output layer (the
        untrained network corrects almost nothing)."""
        rng = np.random.default_rng(seed)
        P = np.zeros(NPARAM, dtype=np.float64)
        o = 0
        for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, NOUT)):
            gain = out_gain if fan_out == NOUT else 1.0
            P[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * gain * np.sqrt(2.0 / (fan_in + fan_out))
            o += fan_in * fan_out + fan_out
        return P.astype(np.float32)

    def forward(self, X, mfma: bool = True):
        """Network outputs c [n, 5] for raw features X [n, 8] (FP32 device)."""
        import torch

        X = X.contiguous()
        out = torch.empty((X.shape[0], NOUT), dtype=torch.float32, device=X.device)
        d = _FwdDesc(int(X.shape[0]), int(bool(mfma)), _ptr(X), _ptr(self.P), _ptr(self.norm), _ptr(out))
        check(lib().bc_closure_forward(ctypes.byref(d), _stream()))
        return out

    def save(self, path):
        np.savez(path, P=self.P.cpu().numpy(), norm=self.norm.cpu().numpy())

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path) as f:
            return cls(f["P"], f["norm"], device=device)


# -------------------------------------------------------------------- data
@dataclasses.dataclass
class ClosureData:
    X: object  # [n, 8] raw features (FP32 device)
    D: object  # [n, 5] plain exp-Euler increments of the gates
    Y: object  # [n, 5] fine end state minus the plain update

    @property
    def n(self):
        return int(self.X.shape[0])

    def stats(self):
        """mu, sd of the features and per-output scales of the targets."""
        mu = self.X.double().mean(0)
        sd = self.X.double().std(0)
        sd = self.X.new_tensor(np.where(sd.cpu().numpy() > 1e-12, sd.cpu().numpy(), 1.0)).double()
        sc = self.Y.double().std(0).clamp_min(1e-12)
        import torch

        return torch.cat([mu, sd, sc]).float()


def generate_data(models, *, dt_fine: float = 0.005, dt_coarse: float = 0.1, n_pairs: int = 100,
                  warm_ms: float = 5.0, gap_ms: float = 0.5, threads: int = 256,
                  features: str = "coarse") -> ClosureData:
    """Training pairs from plain ``dt_fine`` trajectories of ``models``:
    after ``warm_ms``, ``n_pairs`` times: state S0, ``dt_coarse`` of fine
    steps, state S1, then ``gap_ms`` more.  One pair per compartment each
    time (n_pairs x total CVs samples).

    The targets are the fine end gates S1.  With ``features="coarse"`` the
    network's voltage and calcium inputs are those the closure sees when it
    runs: one coarse step (gates skipped) from S0, the same state the coarse
    step hands to ``bc_closure_apply``; with ``"fine"`` they are the fine
    trajectory's V and Ca at the step's end."""
    import torch

    if features not in ("coarse", "fine"):
        raise ValueError("features must be 'coarse' or 'fine'")
    L = StaggerLayout(models, threads=threads, probes="soma")
    if L.chset != CHSET:
        raise ValueError("the closure is trained for the Na/KDR/IL/CaL/CalciumDetailed model")
    rt = StaggerRuntime(L, dt_fine, precision=32)
    co = None
    if features == "coarse":
        co = StaggerRuntime(L, dt_coarse, precision=32)
        co.flags = SKIP_GATES
        tc = co.new_trace(1)
    k = int(round(dt_coarse / dt_fine))
    gap = int(round(gap_ms / dt_fine))
    tr = rt.new_trace(max(k, gap, int(round(warm_ms / dt_fine)), 1))
    w = int(round(warm_ms / dt_fine))
    if w:
        rt.run(w, trace=tr[:w])
    ncv = L.n_cv_total
    ns = len(L.slots)
    X = torch.empty((n_pairs * ncv, NIN), dtype=torch.float32, device=rt.device)
    D = torch.empty((n_pairs * ncv, NOUT), dtype=torch.float32, device=rt.device)
    Y = torch.empty((n_pairs * ncv, NOUT), dtype=torch.float32, device=rt.device)
    for i in range(n_pairs):
        S0 = rt.S.clone()
        if co is not None:
            co.V.copy_(rt.V)
            co.S.copy_(rt.S)
            co.Vbnd.copy_(rt.Vbnd)
            co.step_index = int(round(rt.step_index * dt_fine / dt_coarse))
            co.run(1, trace=tc)
        rt.run(k, trace=tr[:k])
        V1, S1 = rt.V, rt.S
        if co is not None:
            V1 = co.V
            S1 = rt.S.clone()
            S1.view(ns, ncv)[NOUT].copy_(co.S.view(ns, ncv)[NOUT])  # Ca input from the coarse step
        d = _DataDesc(ncv, ncv, float(dt_coarse), ctypes.cast(ctypes.pointer(rt._cc), ctypes.c_void_p), _ptr(V1),
                      _ptr(S0), _ptr(S1), ctypes.c_void_p(X.data_ptr() + i * ncv * NIN * 4),
                      ctypes.c_void_p(D.data_ptr() + i * ncv * NOUT * 4),
                      ctypes.c_void_p(Y.data_ptr() + i * ncv * NOUT * 4))
        check(lib().bc_closure_data(ctypes.byref(d), _stream()))
        if gap:
            rt.run(gap, trace=tr[:gap])
    torch.cuda.synchronize()
    rt.check_status()
    if co is not None:
        co.check_status()
    return ClosureData(X, D, Y)


# ---------------------------------------------------------------- training
def train(data: ClosureData, *, epochs: int = 4, batch: int = 8192, lr: float = 2e-3, seed: int = 0,
          net: ClosureNet = None, log=None) -> ClosureNet:
    """Adam on the one-step loss sum_o (delta_o c_o - y_o)^2 (delta and y
    divided by the per-output scale of y), all in this package's kernels;
    the learning rate decays linearly to 10 % over the run."""
    import torch

    norm = data.stats().to(data.X.device)
    mu, sd, sc = norm[:NIN], norm[NIN:2 * NIN], norm[2 * NIN:]
    Xn = ((data.X - mu) / sd).contiguous()
    Yn = (data.Y / sc).contiguous()
    Dn = (data.D / sc).contiguous()
    norm[2 * NIN:] = 1.0  # c is dimensionless: the output scale stays 1
    if net is None:
        net = ClosureNet(ClosureNet.init_params(seed), norm.cpu().numpy(), device=data.X.device)
    else:
        net.norm.copy_(norm)
    P = net.P
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    max_blocks = (batch + 127) // 128
    partial = torch.empty((max_blocks, NPARAM), dtype=torch.float32, device=P.device)
    loss_part = torch.empty(max_blocks, dtype=torch.float32, device=P.device)
    grad = torch.empty(NPARAM, dtype=torch.float32, device=P.device)
    loss = torch.empty(1, dtype=torch.float32, device=P.device)
    g = torch.Generator(device=P.device)
    g.manual_seed(seed)
    n = data.n
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    step = 0
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device=P.device)
        ep_loss = torch.zeros((), dtype=torch.float64, device=P.device)
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch].contiguous()
            d = _TrainDesc(int(idx.numel()), max_blocks, _ptr(Xn), _ptr(Dn), _ptr(Yn), _ptr(idx), _ptr(P), _ptr(partial),
                           _ptr(loss_part), _ptr(grad), _ptr(loss))
            check(lib().bc_closure_grad(ctypes.byref(d), _stream()))
            step += 1
            rate = lr * (1.0 - 0.9 * (step - 1) / max(1, total - 1))
            a = _AdamDesc(step, rate, 0.9, 0.999, 1e-8, _ptr(P), _ptr(m), _ptr(v), _ptr(grad))
            check(lib().bc_closure_adam(ctypes.byref(a), _stream()))
            ep_loss += loss[0]
        history.append(float(ep_loss.item()) / steps_per_epoch)
        if log:
            log(f"closure epoch {ep}: mean normalised loss {history[-1]:.4e}")
    net.history = history
    return net


def one_step_error(net: ClosureNet, data: ClosureData, mfma: bool = True):
    """Mean squared gate-update error per output without and with the
    closure (teacher-forced, on ``data``)."""
    base = (data.Y.double() ** 2).mean(0)
    corr = data.D.double() * net.forward(data.X, mfma=mfma).double()
    return base.cpu().numpy(), ((data.Y.double() - corr) ** 2).mean(0).cpu().numpy()


# ----------------------------------------------------------------- runtime
class ClosureStepper:
    """Coarse staggered steps with the closure: each step is a one-step
    launch of the staggered kernel that updates V and the Ca pool but skips
    the gates (``BC_STEP_SKIP_GATES``), then ``bc_closure_apply`` performs
    the gate update (plain exp-Euler + network correction) in place."""

    def __init__(self, models, dt: float, net: ClosureNet, *, threads: int = 256, layout=None,
                 flag_tol: float = 0.05):
        self.L = layout or StaggerLayout(models, threads=threads, probes="soma")
        if self.L.chset != CHSET:
            raise ValueError("the closure applies to the Na/KDR/IL/CaL/CalciumDetailed model")
        self.rt = StaggerRuntime(self.L, dt, precision=32)
        self.net = net
        self.dt = float(dt)
        import torch

        # per compartment: 1 once the network over-corrected a gate there
        self.over = torch.zeros(self.L.n_cv_total, dtype=torch.uint8, device=self.rt.device)
        self.flag_tol = float(flag_tol)

    def reset_state(self):
        self.rt.reset_state()
        self.over.zero_()

    def unhealthy(self, v_range=(-150.0, 100.0), gate_tol: float = 1e-3):
        """Per cell: True when, since the last call, the network
        over-corrected a gate of one of its compartments (a corrected gate
        beyond ``flag_tol`` outside [0, 1] before the clip), or now a
        compartment's V is outside ``v_range`` or not finite, a gate is
        outside [-gate_tol, 1 + gate_tol] or the calcium concentration is not
        positive.  One device reduction; clears the over-correction flags."""
        import torch

        L, rt = self.L, self.rt
        if getattr(self, "_cell_of_cv", None) is None:
            cell = np.zeros(L.n_cv_total, dtype=np.int64)
            for i, (c0, c) in enumerate(zip(L.cell_cv0, L.cell_ncv)):
                cell[int(c0):int(c0 + c)] = i
            self._cell_of_cv = torch.as_tensor(cell, device=rt.device)
        V = rt.V
        S = rt.S.view(len(L.slots), L.n_cv_total)
        bad = ~torch.isfinite(V) | (V < v_range[0]) | (V > v_range[1])
        gates = S[:NOUT]
        bad |= (~torch.isfinite(gates) | (gates < -gate_tol) | (gates > 1.0 + gate_tol)).any(0)
        ca = S[NOUT]
        bad |= ~torch.isfinite(ca) | (ca <= 0.0)
        bad |= self.over.bool()
        self.over.zero_()
        per_cell = torch.zeros(len(L.cells), dtype=torch.int32, device=rt.device)
        per_cell.index_add_(0, self._cell_of_cv, bad.to(torch.int32))
        return (per_cell > 0).cpu().numpy()

    def run(self, n_steps: int, trace=None, closure: bool = True):
        rt = self.rt
        ncv = self.L.n_cv_total
        if trace is None:
            trace = rt.new_trace(n_steps)
        if not closure:
            rt.flags = 0
            return rt.run(n_steps, trace=trace, chunk=200)
        cc = ctypes.cast(ctypes.pointer(rt._cc), ctypes.c_void_p)
        d = _ApplyDesc(ncv, ncv, self.dt, cc, _ptr(rt.V), _ptr(rt.S), _ptr(self.net.P), _ptr(self.net.norm),
                       _ptr(self.over), self.flag_tol)
        rt.flags = SKIP_GATES
        try:
            for i in range(n_steps):
                rt.run(1, trace=trace[i:i + 1])
                check(lib().bc_closure_apply(ctypes.byref(d), _stream()))
        finally:
            rt.flags = 0
        return trace


def soma_rmse(coarse, fine, ratio: int):
    """RMSE per cell between a coarse trace [n, cells] and a fine trace
    [n * ratio, cells] sampled at the coarse step ends."""
    f = fine[ratio - 1::ratio]
    n = min(f.shape[0], coarse.shape[0])
    return np.sqrt(np.mean((coarse[:n] - f[:n]) ** 2, axis=0))


def spike_times_ms(col, dt: float, v_th: float = 0.0):
    """Threshold crossings of one trace column (row i = V after step i):
    V_{i-1} < v_th <= V_i is a spike at (i + 1) dt, as the reference's
    ``get_spike`` marks the step (``braincell/_base.py:421-431``)."""
    v = np.asarray(col)
    i = np.nonzero((v[:-1] < v_th) & (v[1:] >= v_th))[0] + 1
    return (i + 1) * dt


def spike_error(coarse, fine, dt_coarse: float, dt_fine: float, v_th: float = 0.0):
    """Per cell: (spike count of the fine trace, largest spike-time error in
    ms; inf when the counts differ)."""
    n_c = coarse.shape[0]
    counts, err = [], []
    for j in range(coarse.shape[1]):
        tc = spike_times_ms(coarse[:, j], dt_coarse, v_th)
        tf = spike_times_ms(fine[: int(round(n_c * dt_coarse / dt_fine)), j], dt_fine, v_th)
        counts.append(len(tf))
        err.append(float(np.max(np.abs(tc - tf))) if len(tc) == len(tf) and len(tf) else
                   (0.0 if len(tc) == len(tf) else np.inf))
    return np.asarray(counts), np.asarray(err)


class ClosureBatch:
    """Closure at a coarse dt where the plain fine method accepts it, plain
    ``fallback_dt`` elsewhere.

    ``calibrate`` judges every cell: closure and plain ``judge_dt`` run
    ``window_ms``; a cell is accepted when its soma RMSE is at most
    ``threshold_mV`` and it has the fine run's spikes, each within
    ``spike_tol_ms``.  ``run`` watches the accepted cells while they run:
    every ``check_every`` coarse steps it reads their state, and a cell
    whose voltage, gates or calcium leave the physical range (or are not
    finite) moves to the plain fallback for the whole run.  A non-finite
    value in a plain fallback run raises ``NativeError``."""

    def __init__(self, models, net: ClosureNet, *, dt: float = 0.1, fallback_dt: float = 0.025,
                 judge_dt: float = 0.005, threshold_mV: float = 0.5, spike_tol_ms: float = None,
                 v_range=(-150.0, 100.0)):
        self.models = list(models)
        self.net = net
        self.dt, self.fallback_dt, self.judge_dt = float(dt), float(fallback_dt), float(judge_dt)
        self.threshold = float(threshold_mV)
        self.spike_tol = float(self.dt if spike_tol_ms is None else spike_tol_ms)
        self.v_range = (float(v_range[0]), float(v_range[1]))
        self.accept = np.ones(len(self.models), dtype=bool)
        self.rmse = self.spike_err = self.spike_count = None
        self.moved = []  # (cell index, simulated ms of the check that moved it)

    def calibrate(self, window_ms: float = 20.0):
        import torch

        n = int(round(window_ms / self.dt))
        st = ClosureStepper(self.models, self.dt, self.net)
        st.reset_state()
        coarse = st.run(n).double().cpu().numpy()
        ratio = int(round(self.dt / self.judge_dt))
        L = st.L
        rt = StaggerRuntime(L, self.judge_dt, precision=32)
        fine = rt.run(n * ratio, chunk=2000).double().cpu().numpy()
        torch.cuda.synchronize()
        rt.check_status()
        cols = [int(L.cell_probe0[i]) for i in range(len(self.models))]
        self.rmse = soma_rmse(coarse[:, cols], fine[:, cols], ratio)
        self.spike_count, self.spike_err = spike_error(coarse[:, cols], fine[:, cols], self.dt, self.judge_dt)
        self.accept = (self.rmse <= self.threshold) & (self.spike_err <= self.spike_tol)
        return self.accept

    def run(self, duration_ms: float, check_every: int = 50):
        """Returns (traces of accepted cells at dt, traces of fallback cells
        at fallback_dt, accepted indices, fallback indices); cells moved
        while running are listed in ``self.moved`` and their traces come
        from the fallback run."""
        import torch

        acc = np.nonzero(self.accept)[0]
        self.moved = []
        out_a = None
        if acc.size:
            st = ClosureStepper([self.models[i] for i in acc], self.dt, self.net)
            st.reset_state()
            n = int(round(duration_ms / self.dt))
            out_a = st.rt.new_trace(n)
            alive = np.ones(acc.size, dtype=bool)
            done = 0
            while done < n:
                k = min(int(check_every), n - done)
                st.run(k, trace=out_a[done:done + k])
                done += k
                bad = st.unhealthy(self.v_range) & alive
                for j in np.nonzero(bad)[0]:
                    self.moved.append((int(acc[j]), round(done * self.dt, 6)))
                alive &= ~bad
                st.rt.status.zero_()  # cells that went non-finite have just been moved
            cols = [int(st.L.cell_probe0[j]) for j in np.nonzero(alive)[0]]
            out_a = out_a[:, cols] if cols else None
            acc = acc[alive]
        rej = np.setdiff1d(np.arange(len(self.models)), acc)
        out_r = None
        if rej.size:
            L = StaggerLayout([self.models[i] for i in rej], probes="soma")
            rt = StaggerRuntime(L, self.fallback_dt, precision=32)
            out_r = rt.run(int(round(duration_ms / self.fallback_dt)), chunk=200)
            torch.cuda.synchronize()
            rt.check_status()
        return out_a, out_r, acc, rej
