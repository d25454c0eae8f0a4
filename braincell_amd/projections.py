"""Projections: an edge set expanded into synaptic contacts on a pool.

Mirrors ``braincell.network.projections`` (``network/projections.py:31-540``).
A ``Projection`` takes the rows of one ``EdgeSet`` (all, or ``edge_index``),
asks a contact method how many contacts each edge makes and on which local
target of the postsynaptic synapse pool, then resolves ``weight`` and
``delay`` per contact and emits one ``Connection``.  Everything here runs on
the host at network build time; the result feeds the same contact tables as
a hand-written ``Connection``.

``per_edge`` and ``by_post`` draw targets with
``numpy.random.default_rng(seed).choice`` in the reference's order (edges in
order for ``per_edge``; postsynaptic cells in first-appearance order for
``by_post``), so a seed gives the reference's contacts exactly.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from .edges import EdgeSet, as_index_array, check_bounds, positive_int

__all__ = ["ProjectionEdgeContext", "ContactTable", "ProjectionContactContext", "ContactMethod", "Projection",
           "per_edge", "by_post", "explicit_contacts"]


@dataclass(frozen=True)
class ProjectionEdgeContext:
    """The selected edges a contact method expands (``projections.py:31-60``)."""

    edge_index: np.ndarray
    edge_pre_index: np.ndarray
    edge_post_index: np.ndarray
    pre_size: int
    post_size: int
    pool_size: int
    synapse: str

    @property
    def n_edge(self) -> int:
        return int(self.edge_index.shape[0])


@dataclass
class ContactTable:
    """Contact c comes from selected edge ``source_edge[c]`` and lands on local
    target ``synapse_index[c]`` of the pool (``projections.py:63-101``)."""

    source_edge: object
    synapse_index: object

    def __post_init__(self) -> None:
        self.source_edge = as_index_array(self.source_edge, name="source_edge")
        self.synapse_index = as_index_array(self.synapse_index, name="synapse_index")
        if self.source_edge.shape != self.synapse_index.shape:
            raise ValueError("ContactTable source_edge and synapse_index must have the same shape; "
                             f"got {self.source_edge.shape!r} and {self.synapse_index.shape!r}.")

    @property
    def n_contact(self) -> int:
        return int(self.source_edge.shape[0])

    def __repr__(self) -> str:
        return f"ContactTable(n_contact={self.n_contact})"

    __str__ = __repr__


@dataclass(frozen=True)
class ProjectionContactContext:
    """What weight / delay callables see: the edges, the contacts and their
    endpoint cells (``projections.py:104-144``)."""

    edge_index: np.ndarray
    edge_pre_index: np.ndarray
    edge_post_index: np.ndarray
    source_edge: np.ndarray
    contact_pre_index: np.ndarray
    contact_post_index: np.ndarray
    synapse_index: np.ndarray
    pre_size: int
    post_size: int
    pool_size: int
    synapse: str

    @property
    def n_edge(self) -> int:
        return int(self.edge_index.shape[0])

    @property
    def n_contact(self) -> int:
        return int(self.source_edge.shape[0])


@dataclass(frozen=True)
class ContactMethod:
    """``builder(ProjectionEdgeContext) -> ContactTable`` (``projections.py:147-176``)."""

    builder: Callable

    def __post_init__(self) -> None:
        if not callable(self.builder):
            raise TypeError(f"ContactMethod builder must be callable, got {type(self.builder).__name__}.")

    def __call__(self, context: ProjectionEdgeContext) -> ContactTable:
        table = self.builder(context)
        if not isinstance(table, ContactTable):
            raise TypeError(f"Projection contact method must return ContactTable, got {type(table).__name__}.")
        return table


def _name(value, what: str) -> str:
    if not isinstance(value, str) or not value:
        raise ValueError(f"Projection {what} must be a non-empty string.")
    return value


def _method(value) -> ContactMethod:
    if value is None:
        return per_edge()
    if isinstance(value, ContactMethod):
        return value
    if callable(value):
        return ContactMethod(value)
    raise TypeError(f"Projection method must be a ContactMethod or callable, got {type(value).__name__}.")


def _per_contact(value, ctx: ProjectionContactContext, what: str):
    """Scalar -> as is; per selected edge -> gathered per contact; per contact
    -> as is; a callable is first evaluated on the contact context
    (``projections.py:353-380``)."""
    if value is None:
        return None
    if callable(value):
        value = value(ctx)
    arr = np.asarray(value)
    if arr.shape == ():
        return value
    if arr.shape == (ctx.n_edge,):
        return arr[ctx.source_edge]
    if arr.shape == (ctx.n_contact,):
        return arr
    raise ValueError(f"Projection {what} must be scalar, shape {(ctx.n_edge,)!r}, "
                     f"or shape {(ctx.n_contact,)!r}; got {arr.shape!r}.")


def _counts(number, ctx: ProjectionEdgeContext) -> np.ndarray:
    """Contacts per selected edge: a scalar, an (n_edge,) integer array or a
    callable returning either (``projections.py:383-396``)."""
    if callable(number):
        number = number(ctx)
    arr = np.asarray(number)
    if arr.shape == ():
        arr = np.full(ctx.n_edge, arr, dtype=arr.dtype)
    if arr.shape != (ctx.n_edge,):
        raise ValueError(f"Projection number must be scalar or shape {(ctx.n_edge,)!r}, got {arr.shape!r}.")
    if not np.issubdtype(arr.dtype, np.integer):
        raise TypeError(f"Projection number must contain integers, got {arr.dtype!r}.")
    if np.any(arr < 0):
        raise ValueError("Projection number must be >= 0.")
    return arr.astype(np.int32)


@dataclass
class Projection:
    """Named projection from an edge set onto a postsynaptic synapse pool
    (``projections.py:179-315``).  ``weight`` / ``delay``: scalar, per
    selected edge, per contact, or a callable of the contact context;
    ``weight=None`` uses the pool's default weight.  Units as in
    ``Connection``: weight in uS (dimensionless for the Markov synapses),
    delay in ms."""

    name: str
    edges: object
    synapse: str
    method: object = None
    edge_index: object = None
    weight: object = None
    delay: object = field(default=0.0)

    def __post_init__(self) -> None:
        self.name = _name(self.name, "name")
        if not isinstance(self.edges, (str, EdgeSet)):
            raise TypeError("Projection edges must be an EdgeSet or edge-set name.")
        if isinstance(self.edges, str):
            self.edges = _name(self.edges, "edges")
        self.synapse = _name(self.synapse, "synapse")
        self.method = _method(self.method)
        if self.edge_index is not None:
            self.edge_index = as_index_array(self.edge_index, name="edge_index")

    @property
    def edge_set_name(self) -> str:
        return self.edges if isinstance(self.edges, str) else self.edges.name

    def __repr__(self) -> str:
        return (f"Projection(name={self.name!r}, edges={self.edge_set_name!r}, synapse={self.synapse!r}, "
                f"edge_index={self.edge_index is not None}, weight={self.weight is not None}, "
                f"delay={self.delay is not None})")

    __str__ = __repr__

    def to_connections(self, edge_set: EdgeSet, *, pre_size: int = None, post_size: int = None,
                       pool_size: int = None) -> tuple:
        """Expand into ``Connection`` declarations (one per projection).  The
        network passes the population sizes and the pool's location count;
        direct calls infer the sizes from the edge set and take one target."""
        from .network import Connection

        sel = (np.arange(edge_set.n_edge, dtype=np.int32) if self.edge_index is None else self.edge_index)
        check_bounds(sel, edge_set.n_edge, f"Projection edge_index for EdgeSet {edge_set.name!r}")
        e_pre, e_post = edge_set.pre_index[sel], edge_set.post_index[sel]

        def size(value, idx, what):
            if value is not None:
                return positive_int(value, f"Projection {what}")
            return int(idx.max()) + 1 if idx.size else 1

        pre_size = size(pre_size, edge_set.pre_index, "pre_size")
        post_size = size(post_size, edge_set.post_index, "post_size")
        pool_size = positive_int(1 if pool_size is None else pool_size, "Projection pool_size")
        ectx = ProjectionEdgeContext(edge_index=sel, edge_pre_index=e_pre, edge_post_index=e_post,
                                     pre_size=pre_size, post_size=post_size, pool_size=pool_size,
                                     synapse=self.synapse)
        table = self.method(ectx)
        check_bounds(table.source_edge, ectx.n_edge, "Projection source_edge")
        check_bounds(table.synapse_index, pool_size, "Projection synapse_index")
        c_pre, c_post = e_pre[table.source_edge], e_post[table.source_edge]
        cctx = ProjectionContactContext(edge_index=sel, edge_pre_index=e_pre, edge_post_index=e_post,
                                        source_edge=table.source_edge, contact_pre_index=c_pre,
                                        contact_post_index=c_post, synapse_index=table.synapse_index,
                                        pre_size=pre_size, post_size=post_size, pool_size=pool_size,
                                        synapse=self.synapse)
        weight = _per_contact(self.weight, cctx, "weight")
        delay = _per_contact(self.delay, cctx, "delay")
        return (Connection(edge_set.pre_population, edge_set.post_population, c_pre, c_post, self.synapse,
                           weight=weight, delay=0.0 if delay is None else delay,
                           synapse_index=table.synapse_index),)


def per_edge(number=1, *, replace: bool = True, seed: int = None) -> ContactMethod:
    """``number`` contacts per selected edge, targets drawn independently per
    edge (``projections.py:430-461``)."""

    def _build(ctx: ProjectionEdgeContext) -> ContactTable:
        counts = _counts(number, ctx)
        src = np.repeat(np.arange(ctx.n_edge, dtype=np.int32), counts)
        if src.size == 0:
            return ContactTable(src, np.zeros(0, dtype=np.int32))
        if not replace and np.any(counts > ctx.pool_size):
            raise ValueError("per_edge(..., replace=False) requires every edge number <= pool_size.")
        rng = np.random.default_rng(seed)
        picks = [rng.choice(ctx.pool_size, size=int(k), replace=replace) for k in counts.tolist() if k > 0]
        return ContactTable(src, np.concatenate(picks))

    return ContactMethod(_build)


def by_post(number=1, *, replace: bool = True, seed: int = None) -> ContactMethod:
    """``number`` contacts per selected edge, targets drawn jointly for all
    contacts on one postsynaptic cell (``projections.py:464-503``)."""

    def _build(ctx: ProjectionEdgeContext) -> ContactTable:
        counts = _counts(number, ctx)
        src = np.repeat(np.arange(ctx.n_edge, dtype=np.int32), counts)
        if src.size == 0:
            return ContactTable(src, np.zeros(0, dtype=np.int32))
        rng = np.random.default_rng(seed)
        tgt = np.empty(src.shape[0], dtype=np.int32)
        post = ctx.edge_post_index[src]
        _, first = np.unique(post, return_index=True)
        for cell in post[np.sort(first)]:
            where = np.flatnonzero(post == cell)
            if not replace and where.size > ctx.pool_size:
                raise ValueError("by_post(..., replace=False) requires total contacts per post <= pool_size; "
                                 f"post_index={int(cell)!r}, demand={int(where.size)!r}, "
                                 f"pool_size={ctx.pool_size!r}.")
            tgt[where] = rng.choice(ctx.pool_size, size=int(where.size), replace=replace)
        return ContactTable(src, tgt)

    return ContactMethod(_build)


def explicit_contacts(source_edge, synapse_index=0) -> ContactMethod:
    """Contacts given outright; a scalar ``synapse_index`` applies to all
    (``projections.py:506-527``)."""

    def _build(ctx: ProjectionEdgeContext) -> ContactTable:
        src = as_index_array(source_edge, name="source_edge")
        tgt = np.asarray(synapse_index)
        if tgt.shape == ():
            tgt = np.full(src.shape, tgt, dtype=tgt.dtype)
        if tgt.shape != src.shape:
            raise ValueError(f"explicit_contacts synapse_index must be scalar or shape {src.shape!r}, "
                             f"got {tgt.shape!r}.")
        return ContactTable(src, tgt)

    return ContactMethod(_build)
