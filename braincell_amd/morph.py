"""Minimal morphology object model: branches of frusta joined into a tree.

Mirrors the declaration surface of ``braincell.morph`` that the staggered
solver consumes (``morph/branch.py:195`` ``Branch.from_lengths``,
``morph/morphology.py:314-348`` ``Morphology.from_root`` and
``morph/morphology.py:1412-1494`` ``Morphology.attach``).  Units are plain
floats: lengths and radii in micrometres.  Branch order is insertion order,
as in the reference (``morph/morphology.py:1537-1543``), and every child is
inserted after its parent.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["Branch", "Morphology"]

BRANCH_TYPES = ("soma", "dendrite", "basal_dendrite", "apical_dendrite", "axon", "custom")


@dataclass
class Branch:
    """One unbranched cable made of frustum segments.

    ``lengths[i]`` is the length of segment ``i``; ``radii_proximal[i]`` and
    ``radii_distal[i]`` its end radii (um).
    """

    lengths: np.ndarray
    radii_proximal: np.ndarray
    radii_distal: np.ndarray
    type: str = "custom"

    def __post_init__(self):
        self.lengths = np.atleast_1d(np.asarray(self.lengths, dtype=np.float64))
        self.radii_proximal = np.atleast_1d(np.asarray(self.radii_proximal, dtype=np.float64))
        self.radii_distal = np.atleast_1d(np.asarray(self.radii_distal, dtype=np.float64))
        n = self.lengths.shape[0]
        if self.radii_proximal.shape != (n,) or self.radii_distal.shape != (n,):
            raise ValueError("radii_proximal/radii_distal must have one entry per segment")
        if np.any(self.radii_proximal <= 0) or np.any(self.radii_distal <= 0):
            raise ValueError("All branch radii must be > 0")
        if float(np.sum(self.lengths)) <= 1e-6:
            raise ValueError("Branch total length must be > 1e-6 um")
        if self.type not in BRANCH_TYPES:
            raise ValueError(f"unknown branch type {self.type!r}; expected one of {BRANCH_TYPES}")

    @classmethod
    def _trusted(cls, lengths, radii_proximal, radii_distal, type):
        """Construct without validation (for generators that guarantee the
        invariants checked in ``__post_init__``)."""
        b = cls.__new__(cls)
        b.lengths, b.radii_proximal, b.radii_distal, b.type = lengths, radii_proximal, radii_distal, type
        return b

    @classmethod
    def from_lengths(cls, *, lengths, radii=None, radii_proximal=None, radii_distal=None, type="custom"):
        """Build a branch from segment lengths and either shared boundary radii
        (``n_segments + 1`` values) or per-segment proximal/distal radii."""
        lengths = np.atleast_1d(np.asarray(lengths, dtype=np.float64))
        if radii is not None:
            if radii_proximal is not None or radii_distal is not None:
                raise TypeError("pass either radii or radii_proximal/radii_distal, not both")
            radii = np.asarray(radii, dtype=np.float64)
            if radii.shape != (lengths.shape[0] + 1,):
                raise ValueError("radii must have n_segments + 1 entries")
            return cls(lengths, radii[:-1], radii[1:], type)
        if radii_proximal is None or radii_distal is None:
            raise TypeError("radii or (radii_proximal, radii_distal) is required")
        return cls(lengths, radii_proximal, radii_distal, type)

    @property
    def total_length(self) -> float:
        return float(np.sum(self.lengths))


@dataclass
class Morphology:
    """A rooted tree of branches.

    ``parent[b]`` is the parent branch index (-1 for the root), ``parent_x[b]``
    the normalised attachment coordinate on the parent and ``child_x[b]`` the
    coordinate of the child that touches the parent (0 or 1).
    """

    branches: list = field(default_factory=list)
    parent: list = field(default_factory=list)
    parent_x: list = field(default_factory=list)
    child_x: list = field(default_factory=list)

    @classmethod
    def from_root(cls, branch: Branch) -> "Morphology":
        return cls([branch], [-1], [0.0], [0.0])

    @classmethod
    def from_swc(cls, path, *, mode: str = "neuron") -> "Morphology":
        """Read an SWC file (``morph/morphology.py`` ``from_swc``; the reader
        is :mod:`braincell_amd.swc`)."""
        from .swc import read_swc

        return read_swc(path, mode=mode)

    def attach(self, parent: int, child: Branch, *, parent_x: float = 1.0, child_x: float = 0.0) -> int:
        """Append ``child`` under branch ``parent``; returns the new branch index."""
        if not (0 <= int(parent) < len(self.branches)):
            raise ValueError(f"parent index {parent} out of range")
        if not (0.0 <= float(parent_x) <= 1.0):
            raise ValueError("parent_x must lie in [0, 1]")
        if float(child_x) not in (0.0, 1.0):
            raise ValueError("child_x must be 0.0 or 1.0")
        self.branches.append(child)
        self.parent.append(int(parent))
        self.parent_x.append(float(parent_x))
        self.child_x.append(float(child_x))
        return len(self.branches) - 1

    @property
    def n_branches(self) -> int:
        return len(self.branches)

    def children(self, b: int) -> list:
        return [i for i, p in enumerate(self.parent) if p == b]

    def branch_types(self) -> list:
        return [br.type for br in self.branches]
