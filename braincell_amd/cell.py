"""Front end: ``Cell`` (paint/place declarations) and ``CellBatch``.

Mirrors the parts of ``braincell.Cell`` (``_multi_compartment/cell.py:163-2708``)
that the staggered step touches: construction with a solver name resolved
through the integrator registry (``cell.py:242, 2771-2776``; reassignment
after ``init_state`` is rejected, ``cell.py:334-337``), ``paint`` / ``place``
(``cell.py:383-400``), ``init_state`` / ``reset_state`` (``cell.py:490-552,
2529-2550``), the update order switch (``ion_channel_update_order``) and
``run(dt, duration)`` (``run.py:54-134``).

A ``Cell`` lowers its declarations into a :class:`CellModel`; a ``CellBatch``
runs many cells (uneven trees) in one packed launch.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .channels import IONS
from .filter import resolve_region
from .mech import (CableProperty, Channel, CurrentClamp, CurrentProbe, Ion, MechanismProbe, NetStim, StateProbe,
                   Synapse)
from .model import CellModel
from .policy import CVPerBranch, CVPolicy
from .quad import get_integrator

__all__ = ["Cell", "CellBatch", "CVPerBranch", "RunResult"]

VOLTAGE_ONLY = 4  # BC_STEP_VOLTAGE_ONLY


@dataclass
class RunResult:
    time: np.ndarray  # ms, step start times (run.py:64-71)
    traces: dict
    # voltage trace name -> per column, the times (entries of ``time``) of
    # the steps at which that CV crossed V_th upwards, V_n < V_th <= V_{n+1}
    # (the steps at which the reference's ``Cell.spike`` is 1)
    spikes: dict = field(default_factory=dict)


def spike_cap(n_steps: int, dt: float) -> int:
    """Spike record per column for a run: 500 Hz over its duration, and no
    more than the ceil(n / 2) upward crossings that n steps can hold."""
    return int(max(1, min((n_steps + 1) // 2, np.ceil(n_steps * float(dt) * 0.5) + 8)))


def _spikes_by_column(rt, step0: int, times: np.ndarray, cols) -> list:
    st = rt.spike_steps()
    return [times[st[c] - step0] for c in cols]


class Cell:
    """A multi-compartment cell integrated by a registry-selected solver."""

    def __init__(self, morpho, *, cv_policy=None, solver="staggered_mi355x", precision: int = 32,
                 V_th: float = 0.0, ion_channel_update_order: str = "family", device="cuda", name=None,
                 pop_size=(1,)):
        self.morpho = morpho
        self.cv_policy = cv_policy or CVPerBranch(1)
        if not isinstance(self.cv_policy, CVPolicy):
            raise TypeError(f"cv_policy must be a CVPolicy, got {type(self.cv_policy).__name__}")
        self._solver = get_integrator(solver)
        self.precision = int(precision)
        self.V_th = float(V_th)
        self.ion_channel_update_order = ion_channel_update_order
        self.device = device
        self.name = name
        self._paint = []
        self._place = []
        self._probes = []
        self._synapses = []
        self._netstims = []
        self._synapse_inputs = {}
        self._runtime = None
        self._initialized = False
        ps = (pop_size,) if isinstance(pop_size, int) else tuple(pop_size)
        if len(ps) != 1 or int(ps[0]) <= 0:
            raise ValueError(f"pop_size must be one positive int, got {pop_size!r}")
        self.pop_size = (int(ps[0]),)

    # ---------------------------------------------------------- declarations
    @property
    def solver(self):
        return self._solver

    @solver.setter
    def solver(self, value):
        if self._initialized:
            raise RuntimeError("Cell.solver cannot be reassigned after init_state().")
        self._solver = get_integrator(value)

    def paint(self, region, *mechanisms) -> "Cell":
        for m in mechanisms:
            if not isinstance(m, (CableProperty, Channel, Ion)):
                raise TypeError(f"paint() takes CableProperty/Channel/Ion, got {type(m).__name__}")
        self._paint.append((region, mechanisms))
        self._runtime = None
        return self

    def place(self, locset, *mechanisms) -> "Cell":
        probes = tuple(m for m in mechanisms if isinstance(m, (StateProbe, MechanismProbe, CurrentProbe)))
        clamps = tuple(m for m in mechanisms if isinstance(m, CurrentClamp))
        synapses = tuple(m for m in mechanisms if isinstance(m, Synapse))
        stims = tuple(m for m in mechanisms if isinstance(m, NetStim))
        for m in mechanisms:
            if not isinstance(m, (CurrentClamp, StateProbe, MechanismProbe, CurrentProbe, Synapse, NetStim)):
                raise TypeError(f"place() takes CurrentClamp, Synapse, NetStim or a probe, got {type(m).__name__}")
        for ns in stims:
            self._netstims.append((locset, ns))
        have = {syn.instance_name for _, syn in self._synapses}
        for syn in synapses:
            if syn.instance_name in have:
                raise ValueError(f"Multiple synapses share the instance name {syn.instance_name!r}.")
            have.add(syn.instance_name)
            self._synapses.append((locset, syn))
        names = {p.name for _, ps in self._probes for p in ps}
        for p in probes:
            if p.name in names:
                raise ValueError(f"Multiple probes share the same name {p.name!r}; probe names must be unique.")
            names.add(p.name)
        if clamps:
            self._place.append((locset, clamps))
        if probes:
            self._probes.append((locset, probes))
        self._runtime = None
        return self

    def bind_synapse_input(self, synapse: str, source, *, weight=1.0, transform=None) -> "Cell":
        """Bind an external drive to the synapse pool ``synapse``
        (``cell.py:403-422``): every step its drive gains
        ``transform(source) * weight`` (``source()`` for a callable),
        broadcast to (pop_size, n_locations), in the pool's payload units.
        Callables are sampled on the host once per step before a run starts,
        so they cannot read the run's own state."""
        self._synapse_inputs.setdefault(str(synapse), []).append((source, weight, transform))
        return self

    def _drive_group(self, L, gid0: int, size: int):
        """(first cell, size, NetStims with their CVs, bindings, pool sizes)
        of ``size`` copies of this cell starting at cell ``gid0`` of layout
        ``L``, for the external drive table."""
        host = L.cells[gid0]
        pools = {name: len(loc) for (name, _, loc) in host.synapses or []}
        stims = [(ns, cv) for ns, cv in (host.netstims or []) if cv >= 0]
        return (gid0, int(size), stims, dict(self._synapse_inputs), pools)

    def _set_drive(self, rt, mode: str, step0: int, n_steps: int, dt: float):
        if getattr(rt, "synapse_tables", None) is None:
            if self._synapse_inputs and not self._synapses:
                raise KeyError("bind_synapse_input: no synapse is placed on the cell.")
            return
        from .synapses import drive_times, set_external_drive

        set_external_drive(rt, [self._drive_group(self._layout, 0, 1)],
                           drive_times(mode, step0, n_steps, dt), step0)

    def lower(self) -> CellModel:
        """Lower paint/place rules to per-branch tables (later rules win)."""
        mo = self.morpho
        nb = mo.n_branches
        cm = np.full(nb, 1.0)
        ra = np.full(nb, 100.0)
        v0 = np.full(nb, -65.0)
        channels = {}
        ca_mask = np.zeros(nb, dtype=bool)
        ca_params = None
        E = {"E_Na": IONS["SodiumFixed"]["E"], "E_K": IONS["PotassiumFixed"]["E"],
             "E_Ca": IONS["CalciumFixed"]["E"]}
        for region, mechs in self._paint:
            mask = resolve_region(region, mo)
            for m in mechs:
                if isinstance(m, CableProperty):
                    cm[mask] = m.membrane_capacitance
                    ra[mask] = m.axial_resistivity
                    v0[mask] = m.resting_potential
                elif isinstance(m, Channel):
                    p = m.resolved()
                    spec = channels.setdefault(m.class_name, {"g_branch": np.zeros(nb), "params": {}})
                    extra = {k: v for k, v in p.items() if k != "g_max"}
                    if spec["params"] and spec["params"] != extra:
                        raise ValueError(f"{m.class_name}: non-density parameters must be the same everywhere")
                    spec["params"] = extra
                    spec["g_branch"][mask] = float(p["g_max"])
                elif isinstance(m, Ion):
                    p = m.resolved()
                    if m.class_name == "CalciumDetailed":
                        if ca_params is not None and ca_params != p:
                            raise ValueError("CalciumDetailed parameters must be the same everywhere")
                        ca_params = p
                        ca_mask |= mask
                    elif m.class_name == "SodiumFixed":
                        E["E_Na"] = float(p["E"])
                    elif m.class_name == "PotassiumFixed":
                        E["E_K"] = float(p["E"])
                    elif m.class_name == "CalciumFixed":
                        E["E_Ca"] = float(p["E"])
        if len(set(v0.tolist())) > 1:
            raise ValueError("resting_potential must be the same on all branches of one cell")
        clamps = []
        for locset, mechs in self._place:
            for (b, x) in locset.locations(mo):
                for m in mechs:
                    for (t0, t1, amp) in m.segments():
                        clamps.append((b, x, t0, t1 - t0, amp))
        synapses = [(syn.instance_name, syn, list(locset.locations(mo))) for locset, syn in self._synapses]
        netstims = [(ns, list(locset.locations(mo))) for locset, ns in self._netstims]
        ca = None
        if ca_params is not None:
            ca = {"mask_branch": ca_mask, **{k: ca_params[k] for k in ("d", "tau", "C_rest", "Co", "Ci0", "temp")}}
        counts = self.cv_policy.counts(mo, cm_branch=cm, ra_branch=ra)
        return CellModel(morph=mo, cv_per_branch=counts, cm=cm, ra=ra, V0=float(v0[0]),
                         channels=channels, ca=ca, clamps=clamps, synapses=synapses, netstims=netstims,
                         **E).validate()

    # ----------------------------------------------------------------- state
    def init_state(self, *, probes=None, threads: int = 256):
        """Build the discretisation, layout and device state (gates start at
        the reference's ``init_state`` value only after ``reset_state``)."""
        from .runtime import StaggerLayout

        self._model = self.lower()
        if probes is None:
            probes = self._probe_cvs() if self._probes else "soma"
        self._layout = StaggerLayout([self._model], threads=threads, probes=probes)
        self._runtime = None
        self._initialized = True
        return self

    def reset_state(self):
        if not self._initialized:
            self.init_state()
        if self._runtime is not None:
            self._runtime.reset_state()
            if getattr(self._runtime, "synapse_buffers", None) is not None:
                self._runtime.synapse_buffers["state"].zero_()

    def _is_cable_be(self) -> bool:
        from .quad import backward_euler_mi355x_step

        return self._solver is backward_euler_mi355x_step

    def _cable_runtime_for(self, dt: float):
        """Runtime of the linearised backward Euler on a passive cell."""
        from .cable import CableLayout, CableRuntime

        if not self._initialized:
            self.init_state()
        rt = self._runtime
        if rt is None or rt.dt != float(dt):
            old = rt
            probes = [self._layout.cells[0].probes]
            rt = CableRuntime(CableLayout([self._model], probes=probes), dt, precision=self.precision,
                              device=self.device)
            if old is not None:
                rt.V.copy_(old.V)
                rt.step_index = old.step_index
            self._runtime = rt
        return rt

    def _runtime_for(self, dt: float):
        from .runtime import StaggerRuntime

        if self._is_cable_be():
            return self._cable_runtime_for(dt)
        if not self._initialized:
            self.init_state()
        rt = self._runtime
        if rt is None or rt.dt != float(dt):
            old = rt
            rt = StaggerRuntime(self._layout, dt, precision=self.precision, device=self.device)
            if self.ion_channel_update_order == "integration":
                rt.flags |= 2  # BC_STEP_INTEGRATION_ORDER
            if self._model.synapses:
                from .synapses import attach_cell_synapses

                attach_cell_synapses(rt, self._layout, dt, self.V_th)
            if old is not None:  # keep the current state across a dt change
                rt.V.copy_(old.V)
                rt.S.copy_(old.S)
                rt.Vbnd.copy_(old.Vbnd)
                if self._model.synapses:
                    rt.synapse_buffers["state"].copy_(old.synapse_buffers["state"])
                rt.step_index = old.step_index
            self._runtime = rt
        return rt

    @property
    def runtime(self):
        return self._runtime

    @property
    def V(self):
        """Membrane voltage per CV (mV), CV order of the discretisation."""
        if self._runtime is None:
            raise RuntimeError("no device state yet: run a step or call Cell.run first")
        return self._runtime.V

    # ------------------------------------------------------------------ run
    def advance(self, t: float, dt: float, n_steps: int = 1, *, voltage_only: bool = False):
        """Native advance used by the registered step functions;
        ``voltage_only`` runs the DHS voltage solve alone (gates and the Ca
        pool keep their values)."""
        rt = self._runtime_for(dt)
        n0 = int(round(float(t) / float(dt)))
        rt.step_index = n0
        if self._is_cable_be():
            if voltage_only:
                raise ValueError("voltage_only applies to the staggered kernel")
            import torch

            cols = torch.as_tensor(rt.L.probe_cv.astype(np.int64), device=rt.V.device)
            before = rt.V[cols].clone() if n_steps == 1 else None
            tr = rt.run(n_steps)
            self.last_trace = tr
            self._cable_last_spike = None if before is None else (before, tr[0].clone())
            return tr
        self._set_drive(rt, "update", n0, int(n_steps), float(dt))
        flags = rt.flags
        if voltage_only:
            rt.flags = flags | VOLTAGE_ONLY
        try:
            tr = rt.run(n_steps)
        finally:
            rt.flags = flags
        rt.step_index = n0 + n_steps
        self.last_trace = tr
        return tr

    def update(self, *args):
        """One step through the configured solver (``cell.py:2348-2387``)."""
        return self.solver(self, *args)

    def run(self, *, dt: float, duration: float, chunk: int = 200) -> RunResult:
        """``run.py:54-134``: T = len(arange(0, duration, dt)) steps starting at
        the cell's current time; returns the probe traces."""
        import torch

        rt = self._runtime_for(dt)
        times = np.arange(0.0, float(duration), float(dt))
        n = int(times.shape[0])
        if n == 0:
            raise ValueError("Cell.run(...) produced no timesteps; ensure duration > 0 and dt > 0.")
        step0 = rt.step_index
        start = step0 * float(dt)
        if self._is_cable_be():
            return self._run_cable(rt, n, start + times)
        self._set_drive(rt, "run", step0, n, float(dt))
        if rt._spk is None or rt._spk.cap < spike_cap(n, dt) or rt._spk.v_th != self.V_th:
            rt.enable_spikes(self.V_th, spike_cap(n, dt))
        else:
            rt.clear_spikes()
        if not self._probes:  # no probes placed: the layout's voltage trace
            tr = rt.run(n, chunk=chunk)
            torch.cuda.synchronize()
            rt.check_status()
            V = tr.double().cpu().numpy()
            return RunResult(time=start + times, traces={"V": V},
                             spikes={"V": _spikes_by_column(rt, step0, start + times, range(V.shape[1]))})
        traces, spike_cols = self._run_probes(rt, n, chunk)
        st = {name: _spikes_by_column(rt, step0, start + times, cols) for name, cols in spike_cols.items()}
        return RunResult(time=start + times, traces=traces, spikes=st)

    def _run_cable(self, rt, n: int, times: np.ndarray) -> RunResult:
        """Cell.run with the passive-cell backward Euler: voltage probes, and
        the upward V_th crossings of each probe column."""
        import torch

        cols = rt.L.probe_cv
        prev = rt.V[torch.as_tensor(cols.astype(np.int64), device=rt.V.device)].double().cpu().numpy()
        tr = rt.run(n)
        torch.cuda.synchronize()
        rt.check_status()
        V = tr.double().cpu().numpy()
        before = np.vstack((prev[None, :], V[:-1]))
        up = (before < self.V_th) & (V >= self.V_th)
        self._cable_last_spike = (torch.as_tensor(before[-1]), torch.as_tensor(V[-1]))
        spikes = [times[np.flatnonzero(up[:, k])] for k in range(V.shape[1])]
        if not self._probes:
            return RunResult(time=times, traces={"V": V}, spikes={"V": spikes})
        points = self._probe_points()
        col = {int(c): k for k, c in enumerate(cols)}
        traces, sp = {}, {}
        for p, cvs in points:
            if not isinstance(p, StateProbe):
                raise NotImplementedError(f"probe {p.name!r}: a passive cell stepped by backward_euler records "
                                          "voltages only")
            idx = [col[int(c)] for c in cvs]
            traces[p.name] = V[:, idx]
            sp[p.name] = [spikes[k] for k in idx]
        return RunResult(time=times, traces={k: traces[k] for k in sorted(traces)},
                         spikes={k: sp[k] for k in sorted(sp)})

    @property
    def spike(self):
        """Per voltage-trace column, whether the last step crossed V_th
        upwards (the reference's ``Cell.spike`` after ``update``)."""
        if self._is_cable_be():
            last = getattr(self, "_cable_last_spike", None)
            if last is None:
                raise RuntimeError("no spike record yet: run the cell first")
            a, b = (x.double().cpu().numpy() for x in last)
            return (a < self.V_th) & (b >= self.V_th)
        rt = self._runtime
        if rt is None or rt._spk is None:
            raise RuntimeError("no spike record yet: run the cell first")
        last = rt.step_index - 1
        return np.asarray([s.size > 0 and int(s[-1]) == last for s in rt.spike_steps()])

    # ---------------------------------------------------------------- probes
    def _probe_points(self):
        """(probe, [CV per location]) for every placed probe."""
        from .layout import discretize_cell
        from .probes import probe_cv

        m = self._model if self._initialized else self.lower()
        ncvb = m.n_cv_branch()
        geom = discretize_cell(m.morph, ncvb, m.per_branch(m.cm), m.per_branch(m.ra))
        out = []
        for locset, probes in self._probes:
            cvs = [probe_cv(geom, ncvb, b, x) for (b, x) in locset.locations(m.morph)]
            for p in probes:
                out.append((p, cvs))
        return out

    def _probe_cvs(self):
        return sorted({c for _, cvs in self._probe_points() for c in cvs})

    def _run_probes(self, rt, n, chunk):
        """Run ``n`` steps and sample every probe after each step
        (``run.py:108-134``: traces keyed by probe name, sorted)."""
        import torch

        from .probes import channel_current, gate_slot, synapse_probe

        points = self._probe_points()
        col = {int(c): k for k, c in enumerate(self._layout.cells[0].probes)}
        syns = {syn.instance_name: syn for _, syn in self._synapses}
        is_syn = lambda p: not isinstance(p, StateProbe) and getattr(p, "mechanism", None) in syns
        need_states = any(not isinstance(p, StateProbe) and not is_syn(p) for p, _ in points)
        need_syn = any(is_syn(p) for p, _ in points)
        L = self._layout
        ns = len(L.slots)
        ncv = L.n_cv_total
        rec = None
        if need_syn:  # the synapse update records A, B of every synapse after each step
            rt.syn_trace = torch.empty((n, 2 * rt.synapse_tables.n_syn), dtype=rt.dtype, device=rt.V.device)
            rt.synapse_desc.n_syn = rt.synapse_tables.n_syn
        try:
            if need_states:
                rec = torch.empty((n + 1) * (ns + 1) * ncv, dtype=rt.dtype, device=rt.V.device)
                tr = rt.run(n, chunk=n, rec=rec)
            else:
                tr = rt.run(n, chunk=chunk)
            torch.cuda.synchronize()
        finally:
            syn_tr = getattr(rt, "syn_trace", None)
            rt.syn_trace = None
            if need_syn:
                rt.synapse_desc.syn_trace = None
        rt.check_status()
        trace = tr.double().cpu().numpy()
        snaps = rec.double().cpu().numpy().reshape(n + 1, ns + 1, ncv)[1:] if need_states else None
        syn_rec = syn_tr.double().cpu().numpy().reshape(n, -1, 2) if need_syn else None
        m = self._model
        cc = L.cc
        out = {}
        spike_cols = {}
        for p, cvs in points:
            cvs = np.asarray(cvs, dtype=np.int64)
            if isinstance(p, StateProbe):
                if any(int(c) not in col for c in cvs):
                    raise ValueError(f"probe {p.name!r}: its CVs are not in the layout's probe set; "
                                     "call init_state() after placing probes")
                spike_cols[p.name] = [col[int(c)] for c in cvs]
                out[p.name] = trace[:, spike_cols[p.name]]
            elif is_syn(p):
                tabs = rt.synapse_tables
                cols = []
                for c in cvs:
                    site = [s_ for (g_, nm, j), s_ in tabs.syn_key.items()
                            if g_ == 0 and nm == p.mechanism and int(tabs.syn_cv[s_]) == int(c)]
                    if not site:
                        raise KeyError(f"Probe {p.name!r} could not find a mechanism named {p.mechanism!r} "
                                       f"at CV {int(c)}.")
                    cols.append(site[0])
                A, B = syn_rec[:, cols, 0], syn_rec[:, cols, 1]
                if isinstance(p, MechanismProbe):
                    out[p.name] = synapse_probe(syns[p.mechanism], p.field, A, B)
                else:
                    if any(int(c) not in col for c in cvs):
                        raise ValueError(f"probe {p.name!r}: its CVs are not in the layout's probe set")
                    V = trace[:, [col[int(c)] for c in cvs]]
                    out[p.name] = synapse_probe(syns[p.mechanism], None, A, B, V)
            elif isinstance(p, MechanismProbe):
                out[p.name] = snaps[:, 1 + gate_slot(L.slots, p.mechanism, p.field)][:, cvs]
            else:
                V = snaps[:, 0][:, cvs]
                st = {key: snaps[:, 1 + s][:, cvs] for s, key in enumerate(L.slots)}
                Ci = st.get(("Ca", 0))
                names = [p.mechanism] if p.mechanism is not None else (
                    [c for c in ("CaL_IS2008", "CaT_HM1992") if c in m.channels] if p.ion == "Ca" else None)
                if names is None:
                    raise ValueError(f"probe {p.name!r}: only ion='Ca' is supported")
                if p.mechanism is not None and p.mechanism not in m.channels:
                    raise KeyError(f"Probe {p.name!r} could not find a mechanism named {p.mechanism!r}.")
                I = np.zeros(V.shape)
                for nm in names:
                    I = I + channel_current(m, nm, cvs, V, st, Ci, cc["E_Ca_fixed"], cc["ca_rtzf"], cc["ca_Co"])
                out[p.name] = I
        return {k: out[k] for k in sorted(out)}, {k: spike_cols[k] for k in sorted(spike_cols)}


class CellBatch:
    """Many cells (uneven morphologies) integrated in one packed launch."""

    def __init__(self, cells_or_models, *, precision: int = 32, threads: int = 256, rows_per_pack=None,
                 probes="soma", device="cuda", V_th: float = 0.0):
        from .runtime import StaggerLayout

        self.V_th = float(V_th)
        for c in cells_or_models:
            if isinstance(c, Cell) and c._synapse_inputs:
                raise NotImplementedError("bound synapse inputs belong to Cell.run or a Network population; a "
                                          "CellBatch takes NetStims only")
        models = [c.lower() if isinstance(c, Cell) else c for c in cells_or_models]
        self.layout = StaggerLayout(models, threads=threads, rows_per_pack=rows_per_pack, probes=probes)
        self.precision = precision
        self.device = device
        self._runtime = None

    def runtime(self, dt: float):
        from .runtime import StaggerRuntime

        if self._runtime is None or self._runtime.dt != float(dt):
            self._runtime = StaggerRuntime(self.layout, dt, precision=self.precision, device=self.device)
            if any(c.synapses for c in self.layout.cells):
                from .synapses import attach_cell_synapses

                attach_cell_synapses(self._runtime, self.layout, dt, self.V_th)
        return self._runtime

    def _drive_groups(self):
        """Per cell with synapses: its NetStims (cells in a batch have no
        bound inputs), for the external drive table."""
        out = []
        for gid, c in enumerate(self.layout.cells):
            if c.synapses:
                pools = {name: len(loc) for (name, _, loc) in c.synapses}
                out.append((gid, 1, [(ns, cv) for ns, cv in (c.netstims or []) if cv >= 0], {}, pools))
        return out

    def run(self, *, dt: float, duration: float, chunk: int = 200, use_graph: bool = True) -> RunResult:
        import torch

        rt = self.runtime(dt)
        n = int(np.arange(0.0, float(duration), float(dt)).shape[0])
        step0 = rt.step_index
        times = (step0 + np.arange(n)) * float(dt)
        if getattr(rt, "synapse_tables", None) is not None:
            from .synapses import drive_times, set_external_drive

            set_external_drive(rt, self._drive_groups(), drive_times("run", step0, n, float(dt)), step0)
        if rt._spk is None or rt._spk.cap < spike_cap(n, dt) or rt._spk.v_th != self.V_th:
            rt.enable_spikes(self.V_th, spike_cap(n, dt))
        else:
            rt.clear_spikes()
        trace = rt.new_trace(n)
        if use_graph and n % chunk == 0:
            g = rt.make_graph(n // chunk, chunk, trace)
            g.launch(rt.stream_handle())
            rt.step_index += n
        else:
            rt.run(n, trace=trace, chunk=chunk)
        torch.cuda.synchronize()
        rt.check_status()
        V = trace.double().cpu().numpy()
        return RunResult(time=times, traces={"V_soma": V},
                         spikes={"V_soma": _spikes_by_column(rt, step0, times, range(V.shape[1]))})

