"""Networks of multi-compartment cell populations: synapses and delayed
spike delivery around the staggered step.

Mirrors ``braincell.network`` (``network/core.py:28-182``,
``engine.py:67-548``, ``delivery.py:255-410``, ``lowering.py:185-228``):
named populations of identical cells (``Cell(pop_size=(n,))``), contact
tables (``Connection``) from a presynaptic population to a named synapse
pool of the postsynaptic cells with a weight (uS) and a delay per contact,
and ``Network.run(dt=..., duration=...)`` returning
``NetworkRunResult(time, traces, spikes)``.

Step semantics (``engine.py:474-548``): at the start of step n each synapse
receives the events that are due, i.e. sent at step n - delay; its states
jump, the voltage solve sees its conductance, then its states decay (the
runtime-synapse phase of the staggered step).  A cell spikes at step n when
any of its CVs crosses V_th upwards during the step; the event arrives at
step n + delay, with delays of ceil(delay / dt) steps and at least one
(``lowering.py:185-228``).

On the GPU all populations share one packed launch.  A launch runs at most
min_delay steps, so every arrival it needs comes from spikes of earlier
launches; the kernel gathers them per synapse, in a fixed order, from a
spike history ring of at least max_delay + launch steps.  No atomics, so
results do not depend on scheduling.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import edges as _edges
from .edges import EdgeMethod, EdgeSet, all_pairs, dense, pairs, probability
from .mech import StateProbe
from .synapses import SynapseTables, drive_times, set_external_drive
from .projections import (ContactMethod, ContactTable, Projection, ProjectionContactContext, ProjectionEdgeContext,
                          by_post, explicit_contacts, per_edge)

__all__ = ["Population", "Connection", "NetworkRunResult", "Network", "delay_steps",
           "EdgeSet", "EdgeMethod", "pairs", "dense", "all_pairs", "probability",
           "Projection", "ContactTable", "ContactMethod", "ProjectionEdgeContext", "ProjectionContactContext",
           "per_edge", "by_post", "explicit_contacts"]

SYN_NPAR = 8  # kernel parameters per synapse (bc_synapse_desc.par)

SPIKE_RECORDINGS = ("full", "population", "none")


def _name(value, what):
    if not isinstance(value, str) or not value:
        raise ValueError(f"{what} must be a non-empty string.")
    return value


def _index_array(value, name):
    arr = np.asarray(value)
    if arr.ndim == 0:
        arr = arr.reshape(1)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be one-dimensional.")
    if arr.size and not np.issubdtype(arr.dtype, np.integer):
        raise TypeError(f"{name} must hold integers.")
    return arr.astype(np.int64)


def delay_steps(delay_ms, dt_ms: float, quantization: str = "ceil") -> np.ndarray:
    """Fixed-step delays (``lowering.py:185-228``): raw = delay / dt snapped to
    the nearest integer within 1e-7, then ceil / floor / strict, at least 1."""
    if quantization not in ("ceil", "floor", "strict"):
        raise ValueError(f"Network delay_quantization must be 'ceil', 'floor', or 'strict', got {quantization!r}.")
    d = np.asarray(delay_ms, dtype=float)
    if np.any(d < 0.0):
        raise ValueError("Connection delay must be >= 0.")
    raw = d / float(dt_ms)
    r = np.rint(raw)
    raw = np.where(np.isclose(raw, r, rtol=1e-7, atol=1e-7), r, raw)
    if quantization == "strict":
        steps = np.rint(raw)
        if not np.allclose(raw, steps, rtol=1e-9, atol=1e-9):
            raise ValueError("Connection delay must be an integer multiple of dt when delay_quantization='strict'.")
    elif quantization == "ceil":
        steps = np.ceil(raw - 1e-12)
    else:
        steps = np.floor(raw + 1e-12)
    return np.maximum(steps.astype(np.int64), 1)


@dataclass
class Population:
    """A named population of identical cells (``core.py:28-80``)."""

    name: str
    cell: object

    def __post_init__(self):
        _name(self.name, "Population name")
        ps = tuple(getattr(self.cell, "pop_size", ()))
        if len(ps) != 1 or int(ps[0]) <= 0:
            raise ValueError(f"A population needs a cell with a one-dimensional positive pop_size, got {ps!r}.")

    @property
    def size(self) -> int:
        return int(self.cell.pop_size[0])

    def __repr__(self) -> str:
        return (f"Population(name={self.name!r}, size={self.size}, cell={type(self.cell).__name__}, "
                f"initialized={bool(getattr(self.cell, '_initialized', False))})")

    __str__ = __repr__


@dataclass
class Connection:
    """Contact table between two populations (``core.py:82-165``): contact c
    sends the spikes of ``pre_index[c]`` to location ``synapse_index[c]`` of
    the synapse pool ``synapse`` on cell ``post_index[c]``.  ``weight`` (uS
    for ExpSyn / Exp2Syn, dimensionless for AMPA / GABAa / NMDA; scalar or
    per contact; None: the pool's default weight, which the Markov types do
    not have) and ``delay`` (ms, scalar or per contact)."""

    pre_population: str
    post_population: str
    pre_index: object
    post_index: object
    synapse: str
    weight: object = None
    delay: object = 0.0
    synapse_index: object = None

    def __post_init__(self):
        _name(self.pre_population, "pre_population")
        _name(self.post_population, "post_population")
        _name(self.synapse, "synapse")
        self.pre_index = _index_array(self.pre_index, "pre_index")
        self.post_index = _index_array(self.post_index, "post_index")
        if self.pre_index.shape != self.post_index.shape:
            raise ValueError("Connection pre_index and post_index must have the same shape; "
                             f"got {self.pre_index.shape!r} and {self.post_index.shape!r}.")
        if self.synapse_index is None:
            self.synapse_index = np.zeros(self.pre_index.shape, dtype=np.int64)
        else:
            self.synapse_index = _index_array(self.synapse_index, "synapse_index")
            if self.synapse_index.shape != self.pre_index.shape:
                raise ValueError("Connection synapse_index must have the same shape as pre_index.")

    @property
    def n_contact(self) -> int:
        return int(self.pre_index.shape[0])

    @property
    def n_edge(self) -> int:
        """Alias of ``n_contact`` (``core.py:109-117``)."""
        return self.n_contact

    def __repr__(self) -> str:
        return (f"Connection(pre_population={self.pre_population!r}, post_population={self.post_population!r}, "
                f"synapse={self.synapse!r}, n_contact={self.n_contact})")

    __str__ = __repr__

    def per_contact(self, value, what):
        arr = np.asarray(value, dtype=float)
        if arr.ndim == 0:
            return np.full(self.n_contact, float(arr))
        if arr.shape != (self.n_contact,):
            raise ValueError(f"Connection {what} must be scalar or shape {(self.n_contact,)!r}, got {arr.shape!r}.")
        return arr


@dataclass(frozen=True)
class NetworkRunResult:
    """``core.py:167-182``: step-start times (ms), traces[population][probe]
    [n_steps, size (, n_locations)] and spikes[population] [n_steps, size]."""

    time: np.ndarray
    traces: dict = field(default_factory=dict)
    spikes: dict = field(default_factory=dict)


class Network:
    """Populations and connections integrated together (``engine.py:67-548``).

    Contacts come from direct ``Connection`` declarations and from
    projections of named edge sets; the run lowers both, direct connections
    first (``engine.py:388``)."""

    def __init__(self, name: str = None):
        if name is not None and (not isinstance(name, str) or not name):
            raise ValueError("Network name must be a non-empty string or None.")
        self.name = name
        self.populations = {}
        self.connections = []
        self.edge_sets = {}
        self.proj = {}
        self.projections = self.proj
        self._run_cache = None  # (run key, tables, device runtime): the state a next run continues from

    def _mark_topology_changed(self) -> None:
        self._run_cache = None

    def __repr__(self):
        return (f"Network(name={self.name!r}, populations={len(self.populations)}, "
                f"edge_sets={len(self.edge_sets)}, projections={len(self.projections)}, "
                f"connections={len(self.connections)})")

    def __str__(self):
        """Multi-line summary (``engine.py:96-136``)."""
        lines = [repr(self), "  populations:"]
        lines += [f"    {n}: size={p.size}, cell={type(p.cell).__name__}" for n, p in self.populations.items()] \
            or ["    <none>"]
        lines.append("  edge_sets:")
        lines += [f"    {n}: {e.pre_population} -> {e.post_population}, n_edge={e.n_edge}"
                  for n, e in self.edge_sets.items()] or ["    <none>"]
        lines.append("  projections:")
        lines += [f"    {n}: edges={p.edge_set_name!r}, synapse={p.synapse!r}" for n, p in self.proj.items()] \
            or ["    <none>"]
        lines.append("  direct connections:")
        lines += [f"    [{i}]: {c.pre_population} -> {c.post_population}, synapse={c.synapse!r}, "
                  f"n_contact={c.n_contact}" for i, c in enumerate(self.connections)] or ["    <none>"]
        return "\n".join(lines)

    def add_population(self, name: str, cell) -> Population:
        if name in self.populations:
            raise ValueError(f"Population {name!r} already exists.")
        pop = Population(name, cell)
        self.populations[name] = pop
        self._mark_topology_changed()
        return pop

    def add_connection(self, connection: Connection) -> Connection:
        if not isinstance(connection, Connection):
            raise TypeError(f"add_connection expects a Connection, got {type(connection).__name__}.")
        self._check_connection(connection)
        self.connections.append(connection)
        self._mark_topology_changed()
        return connection

    def _check_connection(self, connection: Connection) -> None:
        for p in (connection.pre_population, connection.post_population):
            if p not in self.populations:
                raise ValueError(f"Unknown population {p!r}.")
        pre, post = self.populations[connection.pre_population], self.populations[connection.post_population]
        for idx, size, what in ((connection.pre_index, pre.size, "pre_index"),
                                (connection.post_index, post.size, "post_index")):
            if idx.size and (idx.min() < 0 or idx.max() >= size):
                raise ValueError(f"Connection {what} out of range for a population of {size}.")

    # ------------------------------------------------- edge sets, projections
    def add_edge_set(self, edge_set: EdgeSet) -> EdgeSet:
        """Store a reusable cell-level edge set (``engine.py:155-160``)."""
        if not isinstance(edge_set, EdgeSet):
            raise TypeError(f"Network.add_edge_set(...) expects EdgeSet, got {type(edge_set).__name__}.")
        if edge_set.name in self.edge_sets:
            raise ValueError(f"Network already has an edge set named {edge_set.name!r}.")
        self.edge_sets[edge_set.name] = edge_set
        self._mark_topology_changed()
        return edge_set

    def add_edges(self, *, name: str, pre: str, post: str, method) -> EdgeSet:
        """Build an edge set between two populations with an edge method or a
        callable ``method(*, n_pre, n_post)`` and store it (``engine.py:163-178``)."""
        if pre not in self.populations:
            raise KeyError(f"Unknown pre population {pre!r}.")
        if post not in self.populations:
            raise KeyError(f"Unknown post population {post!r}.")
        es = _edges.build(name=name, pre=pre, post=post, method=method, n_pre=self.populations[pre].size,
                          n_post=self.populations[post].size)
        return self.add_edge_set(es)

    def make_edges(self, *, name: str, pre: str, post: str, connector: str, **kwargs) -> EdgeSet:
        """``add_edges`` with the builder named by ``connector``: "pairs"
        (``pairs=``), "all_pairs", "dense" (``adjacency=``) or "probability"
        (``engine.py:212-237``)."""
        extra = dict(kwargs)
        if connector == "pairs":
            method = pairs(extra.pop("pairs"))
        elif connector == "dense":
            method = dense(extra.pop("adjacency"))
        elif connector == "all_pairs":
            method, extra = all_pairs(**extra), {}
        elif connector == "probability":
            method, extra = probability(**extra), {}
        else:
            raise ValueError(f"Unknown edge connector {connector!r}.")
        if extra:
            raise TypeError(f"Unexpected connector arguments: {tuple(extra)!r}.")
        return self.add_edges(name=name, pre=pre, post=post, method=method)

    def add_projection(self, projection: Projection = None, *, name: str = None, edges=None, synapse=None,
                       **kwargs) -> Projection:
        """Store a projection, given as an object or as ``Projection`` keyword
        arguments (``engine.py:181-209``)."""
        if projection is None:
            for value, what in ((name, "name"), (edges, "edges"), (synapse, "synapse")):
                if value is None:
                    raise ValueError(f"Network.add_projection(...) requires {what}.")
            projection = Projection(name=name, edges=edges, synapse=synapse, **kwargs)
        elif any(v is not None for v in (name, edges, synapse)) or kwargs:
            raise TypeError("Network.add_projection(...) accepts either a Projection object "
                            "or direct projection keyword arguments, not both.")
        if not isinstance(projection, Projection):
            raise TypeError(f"Network.add_projection(...) expects Projection, got {type(projection).__name__}.")
        if projection.name in self.proj:
            raise ValueError(f"Network already has a projection named {projection.name!r}.")
        self.proj[projection.name] = projection
        self._mark_topology_changed()
        return projection

    def project(self, *, name: str, edges, synapse, **kwargs) -> Projection:
        """``add_projection`` by keywords (``engine.py:239-241``)."""
        return self.add_projection(name=name, edges=edges, synapse=synapse, **kwargs)

    def _pool_size(self, population: str, pool: str) -> int:
        cell = self.populations[population].cell
        for locset, syn in getattr(cell, "_synapses", []):
            if syn.instance_name == pool:
                return len(list(locset.locations(cell.morpho)))
        raise ValueError(f"Population {population!r} has no synapse pool {pool!r}.")

    def _projection_connections(self) -> tuple:
        """Every projection expanded against its stored edge set, with the
        population sizes and the target pool's location count
        (``engine.py:561-581``)."""
        out = []
        for proj in self.proj.values():
            if proj.edge_set_name not in self.edge_sets:
                raise KeyError(f"Projection {proj.name!r} references unknown EdgeSet {proj.edge_set_name!r}.")
            es = self.edge_sets[proj.edge_set_name]
            for p in (es.pre_population, es.post_population):
                if p not in self.populations:
                    raise ValueError(f"Unknown population {p!r}.")
            cons = proj.to_connections(es, pre_size=self.populations[es.pre_population].size,
                                       post_size=self.populations[es.post_population].size,
                                       pool_size=self._pool_size(es.post_population, proj.synapse))
            for con in cons:
                self._check_connection(con)
            out.extend(cons)
        return tuple(out)

    def all_connections(self) -> tuple:
        """Direct connections, then the projections' (the lowering order)."""
        return tuple(self.connections) + self._projection_connections()

    # ------------------------------------------------------------------ build
    def build(self, dt: float, *, delay_quantization: str = "ceil", chunk: int = 200, threads: int = None,
              rows_per_pack: int = None, autotune: bool = False, full_spikes: bool = False):
        """Lower every population into one packed layout with its synapse and
        contact tables.  Pack plan: ``autotune`` times the candidate plans on
        the device; otherwise ``threads`` / ``rows_per_pack`` or, by default,
        about 2048 packs of 64 threads (networks are many small cells).
        ``full_spikes`` puts every CV in the trace (per-CV spike detection)."""
        if not self.populations:
            raise ValueError("Network.run(...) requires at least one population.")
        from .runtime import StaggerLayout

        names = list(self.populations)
        cells = [self.populations[n].cell for n in names]
        precision = {c.precision for c in cells}
        if len(precision) != 1:
            raise ValueError("every population must use the same precision")
        v_th = {float(c.V_th) for c in cells}
        if len(v_th) != 1:
            raise ValueError("every population must use the same V_th")
        models, probes, probe_points = [], [], {}
        pop_cell0, n_cells = {}, 0
        for n, c in zip(names, cells):
            m = c.lower()
            pts = c._probe_points() if c._probes else []
            pools = {syn.instance_name for _, syn in c._synapses}
            for p, _ in pts:
                if not isinstance(p, StateProbe) and getattr(p, "mechanism", None) not in pools:
                    raise NotImplementedError(f"network runs record voltages and synapse states; {p!r} is "
                                              "neither a StateProbe nor a probe of a synapse pool")
            cvs = sorted({cv for _, cvl in pts for cv in cvl})
            if full_spikes:
                cvs = list(range(int(np.sum(m.n_cv_branch()))))
            probe_points[n] = (pts, cvs)
            size = self.populations[n].size
            pop_cell0[n] = n_cells
            n_cells += size
            models += [m] * size
            probes += [np.asarray(cvs, dtype=np.int64)] * size
        layout = StaggerLayout(models, threads=256, probes_per_model=probes)
        if autotune:
            from .autotune import autotune as _autotune

            # network launches run the synapse kernels, which use pointer jumping only
            layout, _ = _autotune(layout, float(dt), precision=precision.copy().pop(), backsub=("jump",))
        elif threads is not None or rows_per_pack is not None:
            layout = layout.repacked(threads=threads or 256, rows_per_pack=rows_per_pack)
        else:
            nrow_max = max(c.geom.n_point for c in layout.cells)
            total = sum(c.geom.n_point for c in layout.cells)
            rows = max(nrow_max, total // 2048)
            t = 64 if max(c.geom.n_cv for c in layout.cells) <= 512 and rows <= 512 else 256
            layout = layout.repacked(threads=t, rows_per_pack=rows)
        if layout.cv_per_thread > 8:
            layout = layout.repacked(threads=256, cv_per_thread=8, rows_per_pack=layout.rows_per_pack)
        if layout.cv_per_thread > 8:
            raise NotImplementedError("network kernels hold at most 8 CVs per thread (cells up to 2048 CVs)")
        return _NetworkTables(self, names, layout, pop_cell0, n_cells, probe_points, float(dt),
                              delay_quantization, int(chunk), precision.pop(), v_th.pop())

    # -------------------------------------------------------------------- run
    def init_state(self, batch_size=None) -> "Network":
        """Populations are declarations here; device state is built by the
        first run (``engine.py:243-266``)."""
        if batch_size is not None:
            raise NotImplementedError("batched network state is not supported")
        return self

    def reset_state(self, batch_size=None) -> "Network":
        """Reset every cell and synapse to its initial state and the clock to
        t = 0 (``engine.py:268-292``); the next run starts there."""
        self.init_state(batch_size)
        if self._run_cache is not None:
            rt = self._run_cache[2]
            rt.reset_state()
            rt.synapse_buffers["state"].zero_()
        return self

    def run(self, *, dt: float, duration: float, delay_quantization: str = "ceil", event_backend: str = "auto",
            brainevent_backend: str = "jax_raw", spike_recording: str = "full", chunk: int = 200) -> NetworkRunResult:
        """Run for ``duration`` ms at step ``dt`` ms (``engine.py:294-366``),
        continuing from the state the previous run left (cells, synapses and
        clock); events still in flight at the end of a run are dropped, as
        the reference resets its delivery queues before every run.  The event
        backends select JAX implementations there and are validated only."""
        if not (float(dt) > 0.0) or not (float(duration) > 0.0):
            raise ValueError("Network.run(...) needs dt > 0 and duration > 0.")
        if event_backend not in ("auto", "scatter", "brainevent"):
            raise ValueError(f"Network event_backend must be 'auto', 'scatter', or 'brainevent', got {event_backend!r}.")
        if spike_recording not in SPIKE_RECORDINGS:
            raise ValueError(f"Network spike_recording must be 'full', 'population', or 'none', "
                             f"got {spike_recording!r}.")
        if not self.populations:
            raise ValueError("Network.run(...) requires at least one population.")
        n_steps = int(np.arange(0.0, float(duration), float(dt)).shape[0])
        if n_steps == 0:
            raise ValueError("Network.run(...) produced no timesteps; ensure duration > 0 and dt > 0.")
        full = spike_recording == "full"
        key = (float(dt), delay_quantization, full, int(chunk))
        cache = self._run_cache
        if cache is None or cache[0] != key:
            tabs = self.build(dt, delay_quantization=delay_quantization, chunk=chunk, full_spikes=full)
            rt = tabs.runtime(n_steps)
            if cache is not None:  # same topology, another mode or dt: carry the state over
                old = cache[2]
                t0 = old.step_index * old.dt
                step0 = int(round(t0 / float(dt)))
                if abs(step0 * float(dt) - t0) > 1e-9 * max(1.0, abs(t0)):
                    raise ValueError(f"the network's time {t0} ms is not a multiple of dt = {dt} ms")
                rt.V.copy_(old.V)
                rt.S.copy_(old.S)
                rt.Vbnd.copy_(old.Vbnd)
                rt.synapse_buffers["state"].copy_(old.synapse_buffers["state"])
                rt.step_index = step0
            self._run_cache = (key, tabs, rt)
        else:
            tabs, rt = cache[1], cache[2]
        step0 = rt.step_index
        tabs.begin_run(rt, step0, n_steps)
        trace = rt.new_trace(n_steps)
        if full:  # every CV is a trace column and a crossing detector (V_n < V_th <= V_{n+1})
            rt.enable_spikes(tabs.v_th, (n_steps + 1) // 2)
        if tabs.has_syn_probes:
            rt.syn_trace = rt.torch.empty((n_steps, 2 * tabs.n_syn), dtype=rt.dtype, device=rt.device)
            rt.synapse_desc.n_syn = tabs.n_syn
        rt.run(n_steps, trace, chunk=tabs.launch_steps)
        rt.torch.cuda.synchronize()
        rt.check_status()
        return tabs.result(rt, trace, n_steps, spike_recording, step0)


class _NetworkTables:
    """Synapse, contact and spike-history tables of a network layout."""

    def __init__(self, net, names, layout, pop_cell0, n_cells, probe_points, dt, quantization, chunk, precision,
                 v_th):
        self.net, self.names, self.L = net, names, layout
        self.pop_cell0, self.n_cells, self.probe_points = pop_cell0, n_cells, probe_points
        self.dt, self.precision, self.v_th = dt, precision, v_th
        L = layout
        # synapses: one per (cell, pool, location), ordered by (pack, row)
        st = SynapseTables(L, dt)
        self.syn = st
        self.syn_key, self.syn_weight, self.syn_par = st.syn_key, st.syn_weight, st.syn_par
        self.syn_row, self.syn_row_off, self.syn_pack_off = st.syn_row, st.syn_row_off, st.syn_pack_off
        self.n_syn = st.n_syn
        # contacts, grouped by target synapse in connection order
        tgt, pre, dly, w = [], [], [], []
        for con in net.all_connections():
            pool_of = self._pool_cells(con.post_population, con.synapse)
            n = con.n_contact
            post_gid = pop_cell0[con.post_population] + con.post_index
            pre_gid = pop_cell0[con.pre_population] + con.pre_index
            s = np.empty(n, dtype=np.int64)
            for c_ in range(n):
                key = (int(post_gid[c_]), con.synapse, int(con.synapse_index[c_]))
                if key not in self.syn_key:
                    raise ValueError(f"synapse_index {int(con.synapse_index[c_])} out of range for pool "
                                     f"{con.synapse!r} ({pool_of} locations)")
                s[c_] = self.syn_key[key]
            if con.weight is None and np.isnan(self.syn_weight[s]).any():
                raise ValueError(f"Connection weight is None, but synapse pool {con.synapse!r} has no default "
                                 "'weight' parameter.")
            weight = self.syn_weight[s] if con.weight is None else con.per_contact(con.weight, "weight")
            tgt.append(s)
            pre.append(pre_gid)
            dly.append(delay_steps(con.per_contact(con.delay, "delay"), dt, quantization))
            w.append(weight)
        if tgt:
            tgt, pre, dly, w = (np.concatenate(x) for x in (tgt, pre, dly, w))
        else:
            tgt = pre = dly = np.zeros(0, dtype=np.int64)
            w = np.zeros(0)
        order = np.argsort(tgt, kind="stable")
        self.con_pre = pre[order].astype(np.int32)
        self.con_delay = dly[order].astype(np.int32)
        self.con_w = w[order]
        self.con_off = np.searchsorted(tgt[order], np.arange(self.n_syn + 1), side="left").astype(np.int32)
        self.min_delay = int(dly.min()) if dly.size else int(chunk)
        self.max_delay = int(dly.max()) if dly.size else int(chunk)
        self.launch_steps = max(1, min(int(chunk), self.min_delay))
        cv_cell = np.empty(L.n_cv_total, dtype=np.int32)
        for gid in range(len(L.cells)):
            cv_cell[L.cell_cv0[gid]: L.cell_cv0[gid] + L.cell_ncv[gid]] = gid
        self.cv_cell = cv_cell

    def _pool_cells(self, population, pool):
        g0 = self.pop_cell0[population]
        for (name, _, loc) in self.L.cells[g0].synapses or []:
            if name == pool:
                return len(loc)
        raise ValueError(f"Population {population!r} has no synapse pool {pool!r}.")

    def runtime(self, n_steps: int):
        """Device runtime with the synapse tables attached; the spike history
        keeps the whole run (at least max_delay + one launch)."""
        import torch

        from ._native import SynapseDesc
        from .runtime import StaggerRuntime

        rt = StaggerRuntime(self.L, self.dt, precision=self.precision)
        dev, dtype = rt.device, rt.dtype
        t = lambda a, ty: torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=ty)
        nz = lambda a, ty, fill=0: t(a if a.size else np.full(1, fill), ty)
        self.spk_len = max(self.max_delay + self.launch_steps, int(n_steps))
        buf = {
            "pack_off": t(self.syn_pack_off, torch.int32), "row": nz(self.syn_row, torch.int32),
            "row_off": t(self.syn_row_off, torch.int32), "par": nz(self.syn_par.reshape(-1), dtype),
            "state": torch.zeros(max(1, 2 * self.n_syn), dtype=dtype, device=dev),
            "con_off": t(self.con_off, torch.int32), "con_pre": nz(self.con_pre, torch.int32),
            "con_delay": nz(self.con_delay, torch.int32, 1), "con_w": nz(self.con_w, dtype),
            "cv_cell": t(self.cv_cell, torch.int32),
            "spk": torch.zeros(self.spk_len * self.n_cells, dtype=torch.uint8, device=dev),
        }
        d = SynapseDesc()
        d.n_cells, d.spk_len = self.n_cells, self.spk_len
        d.min_delay, d.max_delay = self.launch_steps if not self.con_pre.size else self.min_delay, self.max_delay
        d.v_th = self.v_th
        self.syn.fill_desc(d)
        for k, v in buf.items():
            setattr(d, k, v.data_ptr())
        rt.synapse_desc = d
        rt.synapse_buffers = buf
        rt.synapse_tables = self.syn
        self.begin_run(rt, 0, int(n_steps))
        return rt

    def begin_run(self, rt, step0: int, n_steps: int) -> None:
        """Before a run of ``n_steps`` from global step ``step0``: an empty
        spike history long enough for the run (in-flight events dropped) and
        the run's external drive."""
        need = max(self.max_delay + self.launch_steps, int(n_steps))
        buf = rt.synapse_buffers
        if self.spk_len < need:
            self.spk_len = need
            buf["spk"] = rt.torch.zeros(self.spk_len * self.n_cells, dtype=rt.torch.uint8, device=rt.device)
            rt.synapse_desc.spk_len = self.spk_len
            rt.synapse_desc.spk = buf["spk"].data_ptr()
        else:
            buf["spk"].zero_()
        groups = [self.net.populations[n].cell._drive_group(self.L, self.pop_cell0[n], self.net.populations[n].size)
                  for n in self.names]
        set_external_drive(rt, groups, drive_times("network", int(step0), int(n_steps), self.dt), int(step0))

    @property
    def has_syn_probes(self) -> bool:
        return any(not isinstance(p, StateProbe) for n in self.names for p, _ in self.probe_points[n][0])

    def result(self, rt, trace, n_steps, spike_recording, step0: int = 0) -> NetworkRunResult:
        from .probes import synapse_probe

        tr = trace.double().cpu().numpy()
        L = self.L
        traces, spikes = {}, {}
        rows = (int(step0) + np.arange(n_steps)) % self.spk_len
        spk = rt.synapse_buffers["spk"].view(self.spk_len, self.n_cells).cpu().numpy().astype(bool)[rows]
        syn_rec = None
        if self.has_syn_probes:
            syn_rec = rt.syn_trace.double().cpu().numpy().reshape(n_steps, -1, 2)
            rt.syn_trace = None
            rt.synapse_desc.syn_trace = None
        steps = rt.spike_steps() if spike_recording == "full" else None
        for n in self.names:
            g0, size = self.pop_cell0[n], self.net.populations[n].size
            pts, cvs = self.probe_points[n]
            col = {cv: k for k, cv in enumerate(cvs)}
            traces[n] = {}
            syns = {syn.instance_name: syn for _, syn in self.net.populations[n].cell._synapses}
            for p, cvl in sorted(pts, key=lambda x: x[0].name):
                cols = np.asarray([[L.cell_probe0[g0 + i] + col[cv] for cv in cvl] for i in range(size)])
                v = tr[:, cols]  # [n_steps, size, n_loc]
                if not isinstance(p, StateProbe):
                    sidx = np.empty((size, len(cvl)), dtype=np.int64)
                    for i in range(size):
                        for k, cv in enumerate(cvl):
                            hit = [s_ for j in range(self._pool_cells(n, p.mechanism))
                                   for s_ in [self.syn_key[(g0 + i, p.mechanism, j)]]
                                   if int(self.syn.syn_cv[s_]) == int(cv)]
                            if not hit:
                                raise KeyError(f"Probe {p.name!r} could not find a mechanism named "
                                               f"{p.mechanism!r} at CV {int(cv)}.")
                            sidx[i, k] = hit[0]
                    A, B = syn_rec[:, sidx, 0], syn_rec[:, sidx, 1]
                    field = p.field if hasattr(p, "field") else None
                    v = synapse_probe(syns[p.mechanism], field, A, B, v)
                traces[n][p.name] = v[:, :, 0] if len(cvl) == 1 else v
            if spike_recording == "population":
                spikes[n] = spk[:, g0:g0 + size]
            elif spike_recording == "full":
                full = np.zeros((n_steps, size, len(cvs)), dtype=bool)
                for i in range(size):
                    for k in range(len(cvs)):
                        st = steps[L.cell_probe0[g0 + i] + k]  # global steps of this run
                        full[st - int(step0), i, k] = True
                spikes[n] = full
        return NetworkRunResult(time=(int(step0) + np.arange(n_steps)) * self.dt, traces=traces, spikes=spikes)
