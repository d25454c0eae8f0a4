"""Region and locset selectors used by ``Cell.paint`` / ``Cell.place``.

A minimal counterpart of ``braincell.filter``: ``AllRegion``, ``branch_in``
(by branch attribute) and ``RootLocation`` / ``BranchLocation``.  Regions
resolve to a boolean mask over branches; locsets to ``(branch, x)`` pairs.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["AllRegion", "branch_in", "BranchIn", "RootLocation", "BranchLocation", "resolve_region"]


@dataclass(frozen=True)
class AllRegion:
    def mask(self, morpho) -> np.ndarray:
        return np.ones(morpho.n_branches, dtype=bool)


@dataclass(frozen=True)
class BranchIn:
    attribute: str
    values: tuple

    def mask(self, morpho) -> np.ndarray:
        if self.attribute == "type":
            vals = [b.type for b in morpho.branches]
        elif self.attribute == "index":
            vals = list(range(morpho.n_branches))
        else:
            raise ValueError(f"branch_in: unsupported attribute {self.attribute!r} (use 'type' or 'index')")
        return np.asarray([v in self.values for v in vals], dtype=bool)


def branch_in(attribute: str, values) -> BranchIn:
    if isinstance(values, (str, int)):
        values = (values,)
    return BranchIn(attribute, tuple(values))


@dataclass(frozen=True)
class RootLocation:
    """Location ``x`` on the root branch."""

    x: float = 0.5

    def locations(self, morpho):
        return [(0, float(self.x))]


@dataclass(frozen=True)
class BranchLocation:
    branch: int
    x: float = 0.5

    def locations(self, morpho):
        if not (0 <= self.branch < morpho.n_branches):
            raise ValueError(f"branch {self.branch} out of range")
        return [(int(self.branch), float(self.x))]


_NAMED = {
    "all": AllRegion(),
    "soma": BranchIn("type", ("soma",)),
    "dend": BranchIn("type", ("dendrite", "basal_dendrite", "apical_dendrite")),
    "dendrites": BranchIn("type", ("dendrite", "basal_dendrite", "apical_dendrite")),
    "basal": BranchIn("type", ("basal_dendrite",)),
    "apical": BranchIn("type", ("apical_dendrite",)),
}


def resolve_region(region, morpho) -> np.ndarray:
    if isinstance(region, str):
        if region not in _NAMED:
            raise ValueError(f"unknown region name {region!r}; expected one of {sorted(_NAMED)}")
        region = _NAMED[region]
    if not hasattr(region, "mask"):
        raise TypeError(f"expected a region selector, got {type(region).__name__}")
    return region.mask(morpho)
