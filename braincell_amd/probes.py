"""Probe sampling for :meth:`Cell.run` (``_multi_compartment/probes.py``).

Voltage probes are written by the staggered kernel into its trace.  Gate,
pool and current probes are read from the kernel's per-step state
snapshots (the recording mode the adjoint uses): the value after step n is
snapshot n + 1.  Currents are the reference's ohmic ``g_max * gating *
(E - V)`` in uA/cm^2 (``channel/_base.py:658-674``), evaluated from the
recorded voltage and states.
"""

from __future__ import annotations

import numpy as np

from .channels import CHANNELS
from .mech import CurrentProbe, MechanismProbe, StateProbe

__all__ = ["probe_cv", "gate_slot", "channel_current", "synapse_probe", "CA_CHANNELS"]

CA_CHANNELS = ("CaL_IS2008", "CaT_HM1992")
_GATING = {"Na_Ba2002": (3, 1), "Na_HH1952": (3, 1), "KDR_Ba2002": (4,), "K_HH1952": (4,), "CaL_IS2008": (2, 1),
           "CaT_HM1992": (2, 1), "AHP_De1994": (2,), "HCN_HM1992": (1,), "IL": (), "K_Leak": ()}


def probe_cv(geom, n_cv_branch, branch: int, x: float) -> int:
    """CV sampled at (branch, x): the owning CV's midpoint, or for a
    boundary node the lowest-numbered CV touching it (``probes.py:246-253``)."""
    kind, idx = geom.locate(int(branch), float(x), n_cv_branch)
    if kind == "mid":
        return int(idx)
    node = int(geom.row_to_point_id[idx])
    mids = geom.cv_to_node
    touching = np.nonzero((geom.node_parent[mids] == node) | (mids == geom.node_parent[node]))[0]
    if touching.size == 0:
        raise ValueError(f"point ({branch}, {x}) touches no CV")
    return int(touching.min())


def gate_slot(slots, mechanism: str, field: str) -> int:
    """State-slot index of a probed gate or of the Ca pool."""
    if mechanism == "CalciumDetailed":
        if field not in ("C", "Ci"):
            raise ValueError(f"CalciumDetailed has no field {field!r} (use 'C')")
        key = ("Ca", 0)
    else:
        if mechanism not in CHANNELS:
            raise KeyError(f"no channel named {mechanism!r}")
        names = [g[0] for g in CHANNELS[mechanism].gates]
        if field not in names:
            raise ValueError(f"{mechanism} has no gate {field!r} (gates: {names})")
        key = (mechanism, names.index(field))
    if key not in slots:
        raise KeyError(f"{mechanism!r} is not painted on this cell")
    return slots.index(key)


def channel_current(model, name: str, cv: np.ndarray, V, states: dict, Ci, E_ca_fixed: float, ca_rtzf: float,
                    ca_Co: float):
    """g_max * gating * (E - V) (uA/cm^2) of channel ``name`` at CVs ``cv``;
    ``states[(name, k)]`` are the gate values, ``Ci`` the pool (or None)."""
    spec = model.channels.get(name)
    branch = np.asarray(model_cv_branch(model))[cv]
    if spec is None:
        return np.zeros(np.shape(V))
    g = np.asarray(spec["g_branch"], dtype=np.float64)[branch] * 1e3  # mS/cm^2
    f = np.ones(np.shape(V))
    for k, p in enumerate(_GATING[name]):
        f = f * np.asarray(states[(name, k)], dtype=np.float64) ** p
    params = model.channel_params(name)
    ion = CHANNELS[name].ion
    if ion == "Na":
        E = model.E_Na
    elif ion in ("K", "KCa"):
        E = model.E_K
    elif ion == "Ca":
        E = ca_rtzf * np.log(ca_Co / np.asarray(Ci, dtype=np.float64)) if Ci is not None else E_ca_fixed
    else:
        E = float(params["E"])
    return g * f * (E - np.asarray(V, dtype=np.float64))


def model_cv_branch(model):
    ncv = model.n_cv_branch()
    return np.repeat(np.arange(model.n_branches), ncv)


def is_probe(m) -> bool:
    return isinstance(m, (StateProbe, MechanismProbe, CurrentProbe))


_SYN_FIELDS = {"ExpSyn": ("g",), "Exp2Syn": ("g", "A", "B"), "AMPA": ("g",), "GABAa": ("g",), "NMDA": ("g", "x")}


def synapse_probe(syn, field, A, B, V=None):
    """A synapse probe from the kernel's recorded states after each step
    (``probes.py:136-240`` read the runtime synapse node):

    - MechanismProbe (``field`` given): ExpSyn / Exp2Syn g = B - A (uS;
      ExpSyn keeps A = 0), Exp2Syn A and B; AMPA / GABAa g = B; NMDA g = B,
      x = A (dimensionless);
    - CurrentProbe (``field`` None, voltage ``V`` at the synapse's CV): the
      synapse's ``current(V_post)`` (``synapse/markov.py:115-116, 199-200,
      260-262, 304-306, 355-357``): g (V - e) in nA for ExpSyn / Exp2Syn,
      g_max g (E_rev - V) in uA/cm^2 (mS/cm^2 x mV) for the Markov kinds."""
    kind = syn.synapse_type
    p = syn.resolved()
    g = B if syn.markov else B - A
    if field is None:
        if syn.markov:
            return p["g_max"] * g * (p["E_rev"] - V)
        return g * (V - p["e"])
    if field not in _SYN_FIELDS[kind]:
        raise AttributeError(f"{kind} has no probe field {field!r}; fields: {_SYN_FIELDS[kind]}")
    if field == "g":
        return g
    return A if field in ("A", "x") else B
