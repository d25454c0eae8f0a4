"""Host-side synapse tables shared by network runs and single-cell runs:
the per-synapse records of a packed layout and the external drive table
(NetStim trains and bound inputs).

Record order: one synapse per (cell, pool, location), sorted by (pack,
DHS row), so a pack's synapse rows are contiguous and each row's synapses
follow each other (``bc_synapse_desc.pack_off / row / row_off``).

External drive (``_multi_compartment/cell.py:2426-2487``): before a step
the reference binds to every synapse the sum of three drives: the network's
delayed arrivals (``pre_spike``), the ``NetStim``s placed at the synapse's
point (``_compute/state.py:514-530, 617-626``) and the inputs bound with
``Cell.bind_synapse_input`` (``source() * weight``, after ``transform``).
The last two are known on the host before a run: they form a table
[n_steps, n_ext] with one column per synapse that has any, which the kernel
adds to the step's weighted arrivals.  When they are evaluated depends on
the caller:

- ``Network.run``: at the step's own time, after the arrivals are written
  (``network/engine.py:491-504``);
- ``Cell.run``: once at the run's start time before the first step, then
  after each step at that step's (start) time for the next step
  (``_multi_compartment/run.py:108-134``), so step k of a run sees the drive
  at t_{max(k-1, 0)} and an event due exactly at the run's start arrives
  twice;
- ``Cell.update`` (one step through the registry): after each step for the
  next one (``cell.py:2298-2329``), so step m sees the drive at t_{m-1} and
  the cell's first step none.

A bound callable is evaluated on the host once per step, in step order,
before the run starts: it cannot observe the run's own state (the reference
evaluates it inside the traced step; use a network connection for
cell-to-cell coupling).
"""

from __future__ import annotations

import numpy as np

__all__ = ["SynapseTables", "drive_times", "external_drive", "attach_cell_synapses", "set_external_drive"]


class SynapseTables:
    """Per-synapse records of a packed layout (kernel parameters at ``dt``)."""

    def __init__(self, L, dt: float):
        recs = []  # (pack, row, cell, pool, site, par, default weight, cv)
        for gid, c in enumerate(L.cells):
            for (name, syn, loc) in c.synapses or []:
                w0 = syn.resolved().get("weight", np.nan)  # Markov types: no default weight
                for j, (cv, area, cm) in enumerate(loc):
                    row = int(L.cv_info[L.cell_cv0[gid] + cv] & 0xFFFF)
                    recs.append((int(L.cell_pack[gid]), row, gid, name, j, syn.kernel_par(dt, cm, area), w0,
                                 int(cv)))
        recs.sort(key=lambda r: (r[0], r[1]))
        self.n_syn = len(recs)
        self.syn_key = {(r[2], r[3], r[4]): s for s, r in enumerate(recs)}
        self.syn_cv = np.asarray([r[7] for r in recs], dtype=np.int64)
        self.syn_weight = np.asarray([r[6] for r in recs], dtype=np.float64)
        self.syn_par = np.asarray([r[5] for r in recs], dtype=np.float64).reshape(-1, 8)
        pack = np.asarray([r[0] for r in recs], dtype=np.int64)
        row = np.asarray([r[1] for r in recs], dtype=np.int64)
        new_row = np.ones(len(recs), dtype=bool)
        if len(recs) > 1:
            new_row[1:] = (pack[1:] != pack[:-1]) | (row[1:] != row[:-1])
        starts = np.flatnonzero(new_row)
        self.syn_row = row[starts].astype(np.int32)
        self.syn_row_off = np.concatenate((starts, [len(recs)])).astype(np.int32)
        self.syn_pack_off = np.searchsorted(pack[starts], np.arange(L.n_packs + 1), side="left").astype(np.int32)
        # LDS plan of the kernel: most synapse rows / synapses in one pack
        per_pack = self.syn_row_off[self.syn_pack_off[1:]] - self.syn_row_off[self.syn_pack_off[:-1]]
        self.max_syn = int(per_pack.max()) if self.n_syn else 0
        self.max_syn_rows = int(np.diff(self.syn_pack_off).max()) if self.n_syn else 0

    def fill_desc(self, d) -> None:
        """Sizes of the tables in a bc_synapse_desc."""
        d.n_syn, d.max_syn, d.max_syn_rows = self.n_syn, self.max_syn, self.max_syn_rows

    def device_buffers(self, torch, dev, dtype) -> dict:
        t = lambda a, ty: torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=ty)
        nz = lambda a, ty: t(a if a.size else np.zeros(1), ty)
        return {"pack_off": t(self.syn_pack_off, torch.int32), "row": nz(self.syn_row, torch.int32),
                "row_off": t(self.syn_row_off, torch.int32), "par": nz(self.syn_par.reshape(-1), dtype),
                "state": torch.zeros(max(1, 2 * self.n_syn), dtype=dtype, device=dev)}


def drive_times(mode: str, step0: int, n_steps: int, dt: float) -> np.ndarray:
    """Time (ms) at which the drive of each of ``n_steps`` steps from global
    step ``step0`` is evaluated; NaN where the step gets none."""
    n = np.arange(step0, step0 + n_steps, dtype=np.int64)
    if mode == "network":
        src = n
    elif mode == "run":
        src = np.maximum(n - 1, step0)
    elif mode == "update":
        src = n - 1
    else:
        raise ValueError(f"unknown drive mode {mode!r}")
    t = src.astype(np.float64) * float(dt)
    return np.where(src >= 0, t, np.nan)


def _binding_value(source, weight, transform, shape, pool):
    value = source() if callable(source) else source
    if transform is not None:
        value = transform(value)
    contrib = np.asarray(value, dtype=np.float64) * np.asarray(weight, dtype=np.float64)
    try:
        return np.broadcast_to(contrib, shape)
    except ValueError as exc:
        raise ValueError(f"Bound synapse input for {pool!r} cannot broadcast from shape {contrib.shape!r} "
                         f"to {shape!r}.") from exc


def external_drive(tabs: SynapseTables, groups, times: np.ndarray):
    """External drive table for one run.

    ``groups``: per population, (first cell index in the layout, size,
    netstims [(NetStim, CV)], bindings {pool: [(source, weight, transform)]},
    pool sizes {pool: n_locations}).  ``times``: ``drive_times`` output.
    Returns (ext_slot [n_syn] int32, table [n_steps, n_ext] float64) or
    None when no synapse has an external drive."""
    n_steps = int(times.shape[0])
    valid = ~np.isnan(times)
    cols = {}  # synapse -> column
    parts = []  # (synapse, per-step drive)
    for gid0, size, netstims, bindings, pools in groups:
        per_pool = {}
        for pool, n_loc in pools.items():
            per_pool[pool] = np.zeros((n_steps, size, n_loc))
        used = set()
        # NetStims drive every synapse at their CV (same for every cell)
        for ns, cv in netstims:
            fired = np.where(valid, ns.fired(np.where(valid, times, 0.0)), False)
            if not fired.any():
                continue
            for pool, n_loc in pools.items():
                for j in range(n_loc):
                    s = tabs.syn_key[(gid0, pool, j)]
                    if int(tabs.syn_cv[s]) == int(cv):
                        per_pool[pool][:, :, j] += np.where(fired, float(ns.weight), 0.0)[:, None]
                        used.add(pool)
        for pool, binds in bindings.items():
            if pool not in pools:
                raise KeyError(f"bind_synapse_input: no synapse named {pool!r} is placed on the cell.")
            shape = (size, pools[pool])
            for source, weight, transform in binds:
                if not callable(source) and transform is None:
                    v = _binding_value(source, weight, None, shape, pool)
                    per_pool[pool][valid] += v
                else:
                    for k in np.flatnonzero(valid):
                        per_pool[pool][k] += _binding_value(source, weight, transform, shape, pool)
                used.add(pool)
        for pool in sorted(used):
            for i in range(size):
                for j in range(pools[pool]):
                    s = tabs.syn_key[(gid0 + i, pool, j)]
                    cols[s] = len(parts)
                    parts.append(per_pool[pool][:, i, j])
    if not parts:
        return None
    slot = np.full(max(1, tabs.n_syn), -1, dtype=np.int32)
    for s, c in cols.items():
        slot[s] = c
    return slot, np.stack(parts, axis=1)


def attach_cell_synapses(rt, L, dt: float, v_th: float) -> None:
    """Synapses of cells run on their own: the synapse update with no spike
    ring and no contacts (only the external drive reaches them)."""
    import torch

    from ._native import SynapseDesc

    tabs = SynapseTables(L, dt)
    buf = tabs.device_buffers(torch, rt.device, rt.dtype)
    d = SynapseDesc()
    d.n_cells, d.spk_len, d.min_delay, d.max_delay, d.v_th = len(L.cells), 0, 1, 1, float(v_th)
    tabs.fill_desc(d)
    for k in ("pack_off", "row", "row_off", "par", "state"):
        setattr(d, k, buf[k].data_ptr())
    d.spk = None
    rt.synapse_desc, rt.synapse_buffers, rt.synapse_tables = d, buf, tabs


def set_external_drive(rt, groups, times: np.ndarray, step0: int) -> None:
    """Build and upload the external drive of the next ``len(times)`` steps
    from global step ``step0`` into the runtime's synapse descriptor."""
    import torch

    d = rt.synapse_desc
    out = external_drive(rt.synapse_tables, groups, times)
    if out is None:
        d.ext_slot = d.ext_drive = None
        d.n_ext = d.ext_steps = d.ext_step0 = 0
        rt.synapse_buffers.pop("ext_slot", None)
        rt.synapse_buffers.pop("ext_drive", None)
        return
    slot, table = out
    rt.synapse_buffers["ext_slot"] = torch.as_tensor(slot).to(device=rt.device)
    rt.synapse_buffers["ext_drive"] = torch.as_tensor(np.ascontiguousarray(table)).to(device=rt.device,
                                                                                     dtype=rt.dtype)
    d.n_ext, d.ext_step0, d.ext_steps = int(table.shape[1]), int(step0), int(table.shape[0])
    d.ext_slot = rt.synapse_buffers["ext_slot"].data_ptr()
    d.ext_drive = rt.synapse_buffers["ext_drive"].data_ptr()
