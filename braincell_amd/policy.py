"""Control-volume policies: how many CVs each branch is split into.

Mirrors ``braincell._discretization.policy`` (``policy.py:45-349``): every
policy resolves, per branch, a uniform tiling of normalised branch
coordinates into ``n`` CVs; only the count differs between policies.

- ``CVPerBranch(n)`` (``policy.py:84-122``): the same count everywhere.
- ``CVPerBranchList(counts)`` (``policy.py:124-172``): one count per branch.
- ``MaxCVLen(max_cv_len, keep_odd=True)`` (``policy.py:174-221, 388-396``):
  ceil(L / max_cv_len) CVs (1 if the branch is not longer than the maximum
  by more than 1e-6 um), promoted to the next odd count when ``keep_odd``.
- ``DLambda(d_lambda, frequency=100, keep_odd=True)`` (``policy.py:224-286,
  411-433``): NEURON's d_lambda rule.  lambda_f = 1e5 sqrt(d / (4 pi f Ra
  cm)) um per segment (d = r0 + r1 in um, f in Hz, Ra in ohm cm, cm in
  uF/cm^2); n = ceil(sum L / lambda_f / d_lambda), promoted to odd.
- ``CompositeByTypePolicy(rules, default_policy)`` with
  ``CVPolicyByTypeRule(branch_types, policy)`` (``policy.py:289-385``): the
  last rule naming a branch's type picks its policy.

Lengths are plain floats in micrometres and frequencies in hertz.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .morph import BRANCH_TYPES

__all__ = ["CVPolicy", "CVPerBranch", "CVPerBranchList", "MaxCVLen", "DLambda", "CVPolicyByTypeRule",
           "CompositeByTypePolicy"]

EPS_PARAM = 1e-9
EPS_LEN_UM = 1e-6
D_LAMBDA_UM_FACTOR = 1.0e5


def _uniform(n: int) -> tuple:
    return tuple((float(k) / float(n), float(k + 1) / float(n)) for k in range(n))


def _odd(n: int, keep_odd: bool) -> int:
    return n + 1 if keep_odd and n % 2 == 0 else n


def _check_int(value, what: str) -> int:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{what} must be integer, got {value!r}.")
    if value <= 0:
        raise ValueError(f"{what} must be > 0, got {value!r}.")
    return int(value)


def _positive_float(value, what: str) -> float:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, float, np.integer, np.floating)):
        raise TypeError(f"{what} must be a number, got {value!r}.")
    v = float(value)
    if not np.isfinite(v) or v <= 0.0:
        raise ValueError(f"{what} must be > 0, got {value!r}.")
    return v


class CVPolicy:
    """Base class: ``counts`` gives the CV count of every branch."""

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        raise NotImplementedError

    def resolve_cv_bounds(self, morpho, *, cm_branch=None, ra_branch=None) -> tuple:
        """Branch-wise normalised ``(prox, dist)`` CV intervals."""
        return tuple(_uniform(int(n)) for n in self.counts(morpho, cm_branch=cm_branch, ra_branch=ra_branch))


@dataclass(frozen=True)
class CVPerBranch(CVPolicy):
    cv_per_branch: int = 1

    def __post_init__(self):
        _check_int(self.cv_per_branch, "cv_per_branch")

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        return np.full(morpho.n_branches, int(self.cv_per_branch), dtype=np.int64)


@dataclass(frozen=True)
class CVPerBranchList(CVPolicy):
    cv_per_branch: tuple

    def __post_init__(self):
        try:
            counts = tuple(self.cv_per_branch)
        except TypeError as exc:
            raise TypeError("cv_per_branch must be an iterable of integers.") from exc
        for i, c in enumerate(counts):
            _check_int(c, f"cv_per_branch[{i}]")
        object.__setattr__(self, "cv_per_branch", tuple(int(c) for c in counts))

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        if len(self.cv_per_branch) != morpho.n_branches:
            raise ValueError("cv_per_branch length must match morphology branch count: "
                             f"got {len(self.cv_per_branch)}, expected {morpho.n_branches}.")
        return np.asarray(self.cv_per_branch, dtype=np.int64)


@dataclass(frozen=True)
class MaxCVLen(CVPolicy):
    max_cv_len: float  # um
    keep_odd: bool = True

    def __post_init__(self):
        _positive_float(self.max_cv_len, "max_cv_len")

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        m = float(self.max_cv_len)
        out = np.empty(morpho.n_branches, dtype=np.int64)
        for b, br in enumerate(morpho.branches):
            L = float(np.sum(br.lengths))
            if L <= m + EPS_LEN_UM:
                out[b] = 1
                continue
            n = max(1, int(np.ceil(L / m - EPS_PARAM)))
            out[b] = _odd(n, self.keep_odd)
        return out


@dataclass(frozen=True)
class DLambda(CVPolicy):
    d_lambda: float
    frequency: float = 100.0  # Hz
    keep_odd: bool = True

    def __post_init__(self):
        _positive_float(self.d_lambda, "d_lambda")
        _positive_float(self.frequency, "frequency")

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        nb = morpho.n_branches
        cm = np.full(nb, 1.0) if cm_branch is None else np.broadcast_to(np.asarray(cm_branch, float), (nb,))
        ra = np.full(nb, 100.0) if ra_branch is None else np.broadcast_to(np.asarray(ra_branch, float), (nb,))
        out = np.empty(nb, dtype=np.int64)
        for b, br in enumerate(morpho.branches):
            diam = br.radii_proximal + br.radii_distal
            if np.any(diam <= 0.0):
                raise ValueError("DLambda requires strictly positive branch diameters.")
            lam = D_LAMBDA_UM_FACTOR * np.sqrt(diam / (4.0 * np.pi * float(self.frequency) * ra[b] * cm[b]))
            e = float(np.sum(br.lengths / lam))
            n = max(1, int(np.ceil(e / float(self.d_lambda) - EPS_PARAM)))
            out[b] = _odd(n, self.keep_odd)
        return out


@dataclass(frozen=True)
class CVPolicyByTypeRule:
    branch_types: tuple
    policy: CVPolicy

    def __post_init__(self):
        types = tuple(self.branch_types)
        if not types:
            raise ValueError("CVPolicyByTypeRule.branch_types must be non-empty.")
        for t in types:
            if not isinstance(t, str):
                raise TypeError(f"branch_types entries must be str, got {t!r}.")
            if t not in BRANCH_TYPES:
                raise ValueError(f"unknown branch type {t!r}; expected one of {BRANCH_TYPES}")
        if not isinstance(self.policy, CVPolicy):
            raise TypeError(f"policy must be CVPolicy, got {type(self.policy).__name__}.")
        object.__setattr__(self, "branch_types", types)


@dataclass(frozen=True)
class CompositeByTypePolicy(CVPolicy):
    rules: tuple
    default_policy: CVPolicy

    def __post_init__(self):
        for r in self.rules:
            if not isinstance(r, CVPolicyByTypeRule):
                raise TypeError("CompositeByTypePolicy.rules must contain CVPolicyByTypeRule instances.")
        if not isinstance(self.default_policy, CVPolicy):
            raise TypeError(f"default_policy must be CVPolicy, got {type(self.default_policy).__name__}.")
        object.__setattr__(self, "rules", tuple(self.rules))

    def counts(self, morpho, *, cm_branch=None, ra_branch=None) -> np.ndarray:
        chosen = [self.default_policy] * morpho.n_branches
        for r in self.rules:
            for b, br in enumerate(morpho.branches):
                if br.type in r.branch_types:
                    chosen[b] = r.policy
        cache = {}
        out = np.empty(morpho.n_branches, dtype=np.int64)
        for b, pol in enumerate(chosen):
            if id(pol) not in cache:
                cache[id(pol)] = pol.counts(morpho, cm_branch=cm_branch, ra_branch=ra_branch)
            out[b] = cache[id(pol)][b]
        return out
