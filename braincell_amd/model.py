"""Lowered, solver-facing description of one cell (``CellModel``).

``Cell.paint`` / ``Cell.place`` declarations are lowered into per-branch
tables here, the way the reference lowers paint/place rules into per-CV
mechanism buckets (``_discretization/mechanism.py:514-626``) and dense
runtime layouts with ``g_max`` zeroed where a channel is not painted
(``_compute/bindings.py:1065-1085``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .channels import CHANNELS, IONS

__all__ = ["CellModel"]


@dataclass
class CellModel:
    """Per-branch mechanism tables for one cell.

    Attributes
    ----------
    morph : Morphology
    cv_per_branch : int or sequence of int
    cm, ra : float or per-branch arrays (uF/cm^2, ohm cm)
    V0 : float, initial voltage (mV) used by ``reset_state``
    channels : dict name -> {"g_branch": [n_branch] S/cm^2, "params": dict}
    ca : None or dict with "mask_branch" and CalciumDetailed parameters
    E_Na, E_K, E_Ca : fixed reversal potentials (mV)
    clamps : list of (branch, x, start_ms, duration_ms, amplitude_nA)
    synapses : list of (instance name, mech.Synapse, [(branch, x), ...])
    netstims : list of (mech.NetStim, [(branch, x), ...])
    """

    morph: object
    cv_per_branch: object = 1
    cm: object = 1.0
    ra: object = 100.0
    V0: float = -65.0
    channels: dict = field(default_factory=dict)
    ca: object = None
    E_Na: float = IONS["SodiumFixed"]["E"]
    E_K: float = IONS["PotassiumFixed"]["E"]
    E_Ca: float = IONS["CalciumFixed"]["E"]
    clamps: list = field(default_factory=list)
    synapses: list = field(default_factory=list)
    netstims: list = field(default_factory=list)

    @property
    def n_branches(self) -> int:
        return self.morph.n_branches

    def n_cv_branch(self) -> np.ndarray:
        nb = self.n_branches
        if np.isscalar(self.cv_per_branch):
            return np.full(nb, int(self.cv_per_branch), dtype=np.int64)
        return np.asarray(self.cv_per_branch, dtype=np.int64)

    def per_branch(self, value) -> np.ndarray:
        nb = self.n_branches
        if np.isscalar(value):
            return np.full(nb, float(value))
        arr = np.asarray(value, dtype=np.float64)
        if arr.shape != (nb,):
            raise ValueError("per-branch value has the wrong length")
        return arr

    def channel_params(self, name: str) -> dict:
        out = dict(CHANNELS[name].params)
        out.update(self.channels[name].get("params", {}))
        return out

    def validate(self):
        for name, spec in self.channels.items():
            if name not in CHANNELS:
                raise ValueError(f"unknown channel {name!r}")
            g = np.asarray(spec["g_branch"], dtype=np.float64)
            if g.shape != (self.n_branches,):
                raise ValueError(f"{name}: g_branch must have one value per branch")
            if np.any(g < 0):
                raise ValueError(f"{name}: negative conductance")
            if CHANNELS[name].ion == "KCa" and self.ca is None:
                raise ValueError(f"{name} needs a CalciumDetailed pool")
        return self
