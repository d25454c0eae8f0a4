"""Cell-level adjacency for networks: ``EdgeSet`` and the edge builders.

Mirrors ``braincell.network.edges`` (``network/edges.py:51-384``).  An edge
set names which presynaptic cells talk to which postsynaptic cells, nothing
more: no synapse pool, local target, weight or delay (a ``Projection`` adds
those).  Builders are host-side construction helpers; ``build`` calls one
with the two population sizes, stores the result as int32 index arrays and
checks them against the sizes.

Random builders draw from ``numpy.random.default_rng(seed)`` in the same
order as the reference (one ``rng.random((n_pre, n_post))`` matrix, row-major
``nonzero``), so a seed gives the reference's topology exactly.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

__all__ = ["EdgeSet", "EdgeMethod", "pairs", "dense", "all_pairs", "probability"]


def _nonempty(value, what: str, owner: str) -> str:
    if not isinstance(value, str) or not value:
        raise ValueError(f"{owner} {what} must be a non-empty string.")
    return value


def as_index_array(value, *, name: str) -> np.ndarray:
    """1-D int32 index array from an array-like of integers (``core.py:191-208``)."""
    arr = np.asarray(value)
    if arr.ndim == 0:
        arr = arr.reshape(1)
    if arr.ndim != 1:
        raise ValueError(f"{name} must be one-dimensional, got shape {arr.shape!r}.")
    if arr.size and not np.issubdtype(arr.dtype, np.integer):
        raise TypeError(f"{name} must contain integers, got {arr.dtype!r}.")
    return arr.astype(np.int32)


def check_bounds(indices: np.ndarray, size: int, message: str) -> None:
    """Raise IndexError when any index falls outside [0, size)."""
    if indices.size == 0:
        return
    lo, hi = int(indices.min()), int(indices.max())
    if lo < 0 or hi >= int(size):
        raise IndexError(f"{message} out of range for size {int(size)!r}: min={lo!r}, max={hi!r}.")


def positive_int(value, name: str) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{name} must be an integer, got {type(value).__name__}.")
    if int(value) <= 0:
        raise ValueError(f"{name} must be > 0, got {int(value)!r}.")
    return int(value)


@dataclass
class EdgeSet:
    """Named sparse cell-to-cell adjacency (``edges.py:51-107``): edge e joins
    presynaptic cell ``pre_index[e]`` to postsynaptic cell ``post_index[e]``."""

    name: str
    pre_population: str
    post_population: str
    pre_index: object
    post_index: object

    def __post_init__(self) -> None:
        self.name = _nonempty(self.name, "name", "EdgeSet")
        self.pre_population = _nonempty(self.pre_population, "pre_population", "EdgeSet")
        self.post_population = _nonempty(self.post_population, "post_population", "EdgeSet")
        self.pre_index = as_index_array(self.pre_index, name="pre_index")
        self.post_index = as_index_array(self.post_index, name="post_index")
        if self.pre_index.shape != self.post_index.shape:
            raise ValueError("EdgeSet pre_index and post_index must have the same shape; "
                             f"got {self.pre_index.shape!r} and {self.post_index.shape!r}.")

    @property
    def n_edge(self) -> int:
        return int(self.pre_index.shape[0])

    def __repr__(self) -> str:
        return (f"EdgeSet(name={self.name!r}, pre_population={self.pre_population!r}, "
                f"post_population={self.post_population!r}, n_edge={self.n_edge})")

    __str__ = __repr__


@dataclass(frozen=True)
class EdgeMethod:
    """A builder ``builder(*, n_pre, n_post) -> (pre_index, post_index)``
    (``edges.py:110-137``)."""

    builder: Callable

    def __post_init__(self) -> None:
        if not callable(self.builder):
            raise TypeError(f"EdgeMethod builder must be callable, got {type(self.builder).__name__}.")

    def __call__(self, *, n_pre: int, n_post: int):
        return self.builder(n_pre=n_pre, n_post=n_post)


def build(*, name: str, pre: str, post: str, method, n_pre: int, n_post: int) -> EdgeSet:
    """Run ``method`` for the two population sizes and store the edge set
    (``edges.py:140-186``); indices outside the sizes raise IndexError."""
    n_pre = positive_int(n_pre, "n_pre")
    n_post = positive_int(n_post, "n_post")
    if not callable(method):
        raise TypeError("Network.add_edges(...) method must be an EdgeMethod or callable, "
                        f"got {type(method).__name__}.")
    out = method(n_pre=n_pre, n_post=n_post)
    try:
        pre_index, post_index = out
    except (TypeError, ValueError) as exc:
        raise ValueError("Edge method must return a pair ``(pre_index, post_index)``.") from exc
    es = EdgeSet(name=name, pre_population=pre, post_population=post, pre_index=pre_index, post_index=post_index)
    check_bounds(es.pre_index, n_pre, "EdgeSet pre_index")
    check_bounds(es.post_index, n_post, "EdgeSet post_index")
    return es


def pairs(edge_pairs) -> EdgeMethod:
    """Explicit ``(pre, post)`` rows, shape (n_edge, 2) (``edges.py:240-268``)."""

    def _build(*, n_pre: int, n_post: int):
        arr = np.asarray(edge_pairs)
        if arr.size == 0:
            arr = arr.reshape(0, 2)
        if arr.ndim != 2 or arr.shape[1] != 2:
            raise ValueError("edges.pairs(...) expects shape (n_edge, 2).")
        return arr[:, 0], arr[:, 1]

    return EdgeMethod(_build)


def dense(adjacency) -> EdgeMethod:
    """Nonzero entries of a dense [n_pre, n_post] matrix, row-major
    (``edges.py:271-298``)."""

    def _build(*, n_pre: int, n_post: int):
        arr = np.asarray(adjacency)
        if arr.ndim != 2:
            raise ValueError(f"edges.dense(...) expects a two-dimensional array, got {arr.shape!r}.")
        return np.nonzero(arr)

    return EdgeMethod(_build)


def _selection(value, size: int, name: str) -> np.ndarray:
    if isinstance(value, str):
        if value != "all":
            raise ValueError(f"{name} must be 'all' or integer indices, got {value!r}.")
        return np.arange(size, dtype=np.int32)
    return as_index_array(value, name=name)


def all_pairs(*, pre_indices="all", post_indices="all", allow_self: bool = True) -> EdgeMethod:
    """Cartesian product of the selections, pre-major; ``allow_self=False``
    drops i -> i when the two sizes match (``edges.py:301-337``)."""

    def _build(*, n_pre: int, n_post: int):
        a = _selection(pre_indices, n_pre, "pre_indices")
        b = _selection(post_indices, n_post, "post_indices")
        pre_index = np.repeat(a, b.shape[0])
        post_index = np.tile(b, a.shape[0])
        if not allow_self and n_pre == n_post:
            keep = pre_index != post_index
            pre_index, post_index = pre_index[keep], post_index[keep]
        return pre_index, post_index

    return EdgeMethod(_build)


def probability(*, p: float, seed: int | None = None, allow_self: bool = True) -> EdgeMethod:
    """Independent Bernoulli(p) edges from one uniform draw per (pre, post)
    pair (``edges.py:340-374``)."""
    p = float(p)
    if not 0.0 <= p <= 1.0:
        raise ValueError(f"probability p must be in [0, 1], got {p!r}.")

    def _build(*, n_pre: int, n_post: int):
        keep = np.random.default_rng(seed).random((n_pre, n_post)) < p
        if not allow_self and n_pre == n_post:
            keep[np.arange(n_pre), np.arange(n_pre)] = False
        return np.nonzero(keep)

    return EdgeMethod(_build)
