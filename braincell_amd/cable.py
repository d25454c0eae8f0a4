"""Linearised backward Euler for passive multi-compartment cells.

The reference's ``backward_euler`` (``quad/_backward_euler.py:31-160``)
applied to a ``Cell``: ``apply_standard_solver_step`` (``quad/_util.py:
196-262``) stacks every differential state of the cell along a new last
axis, so it runs only when the voltage V (one value per CV) is the sole
state: a gated channel's states and a calcium pool's concentration live in
point space (``pop_size + (n_point,)``, ``_compute/bindings.py:571-630``),
n_point = C + B + 1, and cannot be stacked with V; nor can a synapse's.
Such cells raise here as they fail there.

For a passive cell (leak channels ``IL`` / ``K_Leak``, fixed ions, current
clamps) one step is, per CV i,

    f_i = I_mem,i(V) / Cm_i - (K V)_i        (cell.py:1892-1999)
    J_i = sum_q df_q / dV_i = -a_i - sum_q K_qi   (jacrev_last_dim, _util.py:265-310)
    V_i <- V_i + dt f_i / (1 - dt J_i)

with a_i = sum g / Cm, I_mem,i / Cm_i = b_i - a_i V_i + clamp_i(t_n) and K
the CV axial operator: the node-tree axial matrix with its CV-boundary rows
eliminated (``build_cv_axial_operator``, ``quad/_staggered.py:393-413``).
A boundary row only couples to CV midpoints, so its elimination adds
-A_ib A_bj / A_bb to K_ij for every pair of its neighbours.  Clamps placed
on a CV boundary point do not enter the CV-space derivative
(``currents.py:166-193`` keeps midpoint clamps only).

Device work: ``bc_cable_be_run`` (one workgroup per cell, voltages in LDS).
"""

from __future__ import annotations

import ctypes

import numpy as np

from .layout import discretize_cell
from .runtime import INT32_MAX, first_step_at

__all__ = ["PASSIVE_CHANNELS", "check_passive", "CableLayout", "CableRuntime"]

PASSIVE_CHANNELS = ("IL", "K_Leak")


def check_passive(model) -> None:
    """Raise like the reference's state stacking for cells it cannot step."""
    gated = sorted(set(model.channels) - set(PASSIVE_CHANNELS))
    others = []
    if gated:
        others.append(f"channel states of {gated}")
    if model.ca is not None:
        others.append("the CalciumDetailed concentration")
    if model.synapses:
        others.append(f"synapse states of {[s[0] for s in model.synapses]}")
    if others:
        raise ValueError("backward_euler on a Cell stacks the differential states along the last axis "
                         "(quad/_util.py:116-120) and needs them all to have V's shape (one value per CV); "
                         f"{' and '.join(others)} live in point space. Only passive cells (IL, K_Leak, fixed "
                         "ions, clamps) can be stepped; use the staggered solver for active cells.")


def cv_operator(geom):
    """K in CSR form (rows and columns in CV order) and its column sums."""
    C, P = geom.n_cv, geom.n_point
    row_cv = np.full(P, -1, dtype=np.int64)
    row_cv[geom.dynamic_rows] = np.arange(C)
    entries = {}

    def add(i, j, v):
        entries[(i, j)] = entries.get((i, j), 0.0) + v

    # the row-scaled tree matrix: A[r, parent] = lowers[r], A[parent, r] = uppers[r], A[r, r] = diag_axial[r]
    nbr = [[] for _ in range(P)]  # (neighbour row, A[row, neighbour])
    for r in range(P):
        p = int(geom.parent_rows[r])
        if p >= 0:
            nbr[r].append((p, float(geom.lowers[r])))
            nbr[p].append((r, float(geom.uppers[r])))
    for r in range(P):
        i = int(row_cv[r])
        if i < 0:
            continue
        add(i, i, float(geom.diag_axial[r]))
        for q, v in nbr[r]:
            if row_cv[q] >= 0:
                add(i, int(row_cv[q]), v)
    for r in range(P):  # eliminate each boundary row: K_ij -= A_ib A_bj / A_bb
        if row_cv[r] >= 0:
            continue
        abb = float(geom.diag_axial[r])
        if any(row_cv[q] < 0 for q, _ in nbr[r]):
            raise NotImplementedError("two adjacent CV boundary points (a branch without CVs)")
        if abb == 0.0:
            continue
        for qi, a_bi in nbr[r]:
            a_ib = next(v for q, v in nbr[qi] if q == r)
            for qj, a_bj in nbr[r]:
                add(int(row_cv[qi]), int(row_cv[qj]), -a_ib * a_bj / abb)
    keys = sorted(entries)
    rows = np.asarray([k[0] for k in keys], dtype=np.int64)
    cols = np.asarray([k[1] for k in keys], dtype=np.int32)
    vals = np.asarray([entries[k] for k in keys], dtype=np.float64)
    off = np.searchsorted(rows, np.arange(C + 1), side="left").astype(np.int64)
    colsum = np.zeros(C)
    np.add.at(colsum, cols, vals)
    return off, cols, vals, colsum


class CableLayout:
    """Host tables of a batch of passive cells for ``bc_cable_be_run``."""

    def __init__(self, models, probes="soma"):
        self.models = list(models)
        if not self.models:
            raise ValueError("need at least one cell")
        cv0, k_off, k_col, k_val, a, b, jac, V0 = [0], [0], [], [], [], [], [], []
        self.clamps = []  # (cell, cv, t0, t1, amp mV/ms)
        probe_cv, probe_off = [], [0]
        self.geoms = []
        for ci, m in enumerate(self.models):
            check_passive(m)
            ncvb = m.n_cv_branch()
            geom = discretize_cell(m.morph, ncvb, m.per_branch(m.cm), m.per_branch(m.ra))
            self.geoms.append(geom)
            C = geom.n_cv
            off, cols, vals, colsum = cv_operator(geom)
            k_off.extend((k_off[-1] + off[1:]).tolist())
            k_col.append(cols)
            k_val.append(vals)
            ai = np.zeros(C)
            bi = np.zeros(C)
            for name, spec in m.channels.items():
                g = np.asarray(spec["g_branch"], dtype=np.float64)[geom.cv_branch] * 1e3  # mS/cm^2
                E = m.channel_params(name).get("E", m.E_K) if name == "IL" else m.E_K
                ai += g / geom.cv_cm
                bi += g * E / geom.cv_cm
            a.append(ai)
            b.append(bi)
            jac.append(-ai - colsum)
            V0.append(np.full(C, float(m.V0)))
            for (br, x, t0, dur, amp) in m.clamps:
                kind, idx = geom.locate(int(br), float(x), ncvb)
                if kind == "mid":  # uA/cm^2 / (uF/cm^2) = mV/ms
                    self.clamps.append((ci, int(idx), float(t0), float(t0) + float(dur),
                                        float(amp) * 1e-3 / (geom.cv_area_um2[idx] * 1e-8) / geom.cv_cm[idx]))
            if isinstance(probes, str):
                pr = {"soma": [geom.locate(0, 0.5, ncvb)[1] if geom.locate(0, 0.5, ncvb)[0] == "mid" else 0],
                      "all": list(range(C)), "none": []}[probes]
            else:
                pr = list(probes[ci]) if isinstance(probes[0], (list, tuple, np.ndarray)) else list(probes)
            probe_cv.extend(int(v) for v in pr)
            probe_off.append(len(probe_cv))
            cv0.append(cv0[-1] + C)
        self.cell_cv0 = np.asarray(cv0, dtype=np.int32)
        self.k_off = np.asarray(k_off, dtype=np.int32)
        self.k_col = np.concatenate(k_col).astype(np.int32)
        self.k_val = np.concatenate(k_val)
        self.a, self.b, self.jac = np.concatenate(a), np.concatenate(b), np.concatenate(jac)
        self.V0 = np.concatenate(V0)
        self.probe_cv = np.asarray(probe_cv, dtype=np.int32)
        self.probe_off = np.asarray(probe_off, dtype=np.int32)
        self.n_probes = int(self.probe_cv.size)
        self.n_cv_total = int(self.cell_cv0[-1])
        self.max_cv = int(np.diff(self.cell_cv0).max())
        self.max_nnz = int(np.diff(self.k_off[self.cell_cv0]).max())
        counts = np.bincount([c[0] for c in self.clamps], minlength=len(self.models)) if self.clamps else \
            np.zeros(len(self.models), dtype=np.int64)
        self.clamp_off = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
        self.clamps.sort(key=lambda c: c[0])
        self.validate()

    def validate(self) -> None:
        """Every index a launch reads lies inside its cell (checked before any
        device work)."""
        n_cell = np.diff(self.cell_cv0)
        if np.any(n_cell <= 0) or self.k_off[-1] != self.k_col.size or np.any(np.diff(self.k_off) < 0):
            raise ValueError("inconsistent CV operator offsets")
        cell_of_cv = np.repeat(np.arange(len(self.models)), n_cell)
        cell_of_nz = np.repeat(cell_of_cv, np.diff(self.k_off))
        if np.any(self.k_col < 0) or np.any(self.k_col >= n_cell[cell_of_nz]):
            raise ValueError("CV operator column outside its cell")
        cell_of_probe = np.repeat(np.arange(len(self.models)), np.diff(self.probe_off))
        if np.any(self.probe_cv < 0) or np.any(self.probe_cv >= n_cell[cell_of_probe]):
            raise ValueError("probe outside its cell")
        for (ci, cv, *_rest) in self.clamps:
            if not 0 <= cv < n_cell[ci]:
                raise ValueError("clamp outside its cell")
        if self.lds_bytes(64) > 160 * 1024:
            raise ValueError(f"a cell of {self.max_cv} CVs and {self.max_nnz} operator entries does not fit the "
                             "LDS in FP64")

    def lds_bytes(self, precision: int = 32) -> int:
        """Dynamic LDS of one workgroup (the kernel's cable_lds_bytes): V
        twice, a, b, 1 - dt J and the CSR values in the working precision,
        row offsets and columns as int32."""
        rs = 8 if precision == 64 else 4
        return (5 * self.max_cv + self.max_nnz) * rs + (self.max_cv + 1 + self.max_nnz) * 4


class CableRuntime:
    """Device state and launches of a ``CableLayout``."""

    def __init__(self, layout: CableLayout, dt: float, precision: int = 32, device="cuda", threads: int = 256):
        import torch

        self.torch, self.L, self.dt = torch, layout, float(dt)
        self.precision = int(precision)
        self.dtype = torch.float32 if self.precision == 32 else torch.float64
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("backward_euler on a Cell runs on the GPU (bc_cable_be_run)")
        self.threads = int(threads)
        L = layout
        t = lambda x, ty: torch.as_tensor(np.ascontiguousarray(x)).to(device=self.device, dtype=ty)
        nz = lambda x, ty, fill=0: t(x if np.size(x) else np.full(1, fill), ty)
        self.buf = {
            "cell_cv0": t(L.cell_cv0, torch.int32), "k_off": t(L.k_off, torch.int32),
            "k_col": nz(L.k_col, torch.int32), "k_val": nz(L.k_val, self.dtype), "a": t(L.a, self.dtype),
            "b": t(L.b, self.dtype), "jac": t(L.jac, self.dtype), "clamp_off": t(L.clamp_off, torch.int32),
            "clamp_cv": nz(np.asarray([c[1] for c in L.clamps], dtype=np.int32), torch.int32),
            "clamp_on": nz(np.asarray([min(first_step_at(c[2], dt), INT32_MAX) for c in L.clamps]), torch.int32),
            "clamp_end": nz(np.asarray([min(first_step_at(c[3], dt), INT32_MAX) for c in L.clamps]), torch.int32),
            "clamp_amp": nz(np.asarray([c[4] for c in L.clamps]), self.dtype),
            "probe_off": t(L.probe_off, torch.int32), "probe_cv": nz(L.probe_cv, torch.int32),
            "probe_col": nz(np.arange(L.n_probes, dtype=np.int32), torch.int32),
        }
        self.V = t(L.V0, self.dtype)
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.step_index = 0

    def reset_state(self, V0=None):
        self.V.copy_(self.torch.as_tensor(self.L.V0 if V0 is None else V0, dtype=self.dtype))
        self.step_index = 0
        self.status.zero_()

    def new_trace(self, n_steps: int):
        return self.torch.empty((n_steps, self.L.n_probes), dtype=self.dtype, device=self.device)

    def run(self, n_steps: int, trace=None):
        """Advance ``n_steps`` steps from ``step_index``; returns the probe
        trace [n_steps, n_probes] (V after each step)."""
        from ._native import CableDesc, check, lib

        if trace is None:
            trace = self.new_trace(n_steps)
        if trace.dtype != self.dtype or trace.shape[-1] != self.L.n_probes or trace.shape[0] < n_steps:
            raise ValueError("trace must be [>= n_steps, n_probes] in the runtime dtype")
        d = CableDesc()
        d.precision, d.n_cells, d.threads, d.max_cv = self.precision, len(self.L.models), self.threads, self.L.max_cv
        for k, v in self.buf.items():
            setattr(d, k, v.data_ptr())
        d.trace_ld = self.L.n_probes
        d.trace = trace.data_ptr() if self.L.n_probes else None
        d.V, d.status = self.V.data_ptr(), self.status.data_ptr()
        d.step0, d.n_steps, d.dt = int(self.step_index), int(n_steps), self.dt
        d.max_nnz = self.L.max_nnz
        stream = ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)
        check(lib().bc_cable_be_run(ctypes.byref(d), stream))
        self.step_index += int(n_steps)
        return trace

    def check_status(self):
        st = int(self.status.item())
        if st != 0:
            from ._native import NativeError

            if st & 2:
                raise NativeError(2, "a cell exceeds the launch's LDS plan (max_cv / max_nnz)")
            raise NativeError(4, "non-finite membrane voltage detected")
