"""Seeded synthetic workloads (morphologies, densities, clamps).

All trees are grown with ``numpy.random.default_rng(seed)``; the draw order
below is part of each workload's definition.

- ``single_tree_cell``: one cell, soma (L = 20 um, r = 10 um) plus 63
  dendrites grown as a random binary tree: one dendrite on the soma's distal
  end, then 31 uniformly chosen leaves split in two.  Each dendrite has
  L ~ U(20, 200) um and a single frustum tapering from diameter U(1, 2) um to
  U(0.5, 1) um.  4 CVs per branch (256 CVs, 321 DHS rows).  Cm 1 uF/cm^2,
  Ra 100 ohm cm, leak 0.0003 S/cm^2 at -70 mV everywhere, Na_Ba2002 0.12 and
  KDR_Ba2002 0.036 S/cm^2 on the soma, 0.2 nA clamp at soma x = 0.5 from 10
  to 60 ms, V0 = -70 mV.
- ``uneven_tree_cell(i)``: seed = i.  Target CV count C ~ U{200..2000},
  B = max(3, C // 4) branches of 4 CVs.  The soma gets one apical and one
  basal dendrite; then until B branches exist a uniformly chosen leaf either
  splits in two (probability 0.5, if two slots remain) or is extended by one
  branch.  Lengths and diameters as above.  Na/KDR/leak everywhere, CaL
  0.002 S/cm^2 plus a CalciumDetailed pool on the dendrites; each cell's four
  densities are scaled by exp(0.3 z), z ~ N(0, 1); soma clamp U(0.1, 0.5) nA
  at x = 0.5 from 10 ms to the end of the run; V0 = -70 mV.
- ``region_tree_cell(i)``: seed = 1_000_000 + i, B = 250 (1000 CVs), the
  same growth; Na/KDR/leak/CaL densities per region (soma, basal, apical),
  i.e. 12 parameters per cell, and a CalciumDetailed pool everywhere.
- ``htc_population``: the thalamocortical single-compartment population.
- ``n144_hh_cell(amp)``: the reference's simulator-comparison cell
  (``examples/profiling/simulator_compare/common.py:30-56``,
  ``backend_braincell.py:62-113``): ``data/n144.swc`` (217 branches, 4 CVs
  each = 868 CVs), classic Hodgkin-Huxley (Na_HH1952 120, K_HH1952 36,
  leak 0.3 mS/cm^2 at -54.3 mV; E_Na 50, E_K -77 mV) at 6.3 degC with
  T = T_ref, Cm 1 uF/cm^2, Ra 1000 ohm cm, V0 -62 mV, a clamp of ``amp``
  nA at branch 0, x = 0.85, from 3 to 5 ms.
"""

from __future__ import annotations

import os

import numpy as np

from .channels import celsius2kelvin
from .model import CellModel
from .morph import Branch, Morphology

__all__ = [
    "single_tree_cell", "uneven_tree_cell", "region_tree_cell", "uneven_tree_batch",
    "htc_population", "HTC_CHANNELS", "BASE_DENSITIES", "N144_SWC", "n144_hh_cell", "n144_amplitudes",
    "N144_PROBES",
]

N144_SWC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "n144.swc")
N144_PROBES = ((0, 0.85), (34, 0.85), (54, 0.85))  # (branch, x) of the recorded voltages

BASE_DENSITIES = {"Na_Ba2002": 0.12, "KDR_Ba2002": 0.036, "IL": 0.0003, "CaL_IS2008": 0.002}
CLAMP_FOREVER = 1.0e9


def _dendrite(rng, type_):
    L = rng.uniform(20.0, 200.0)
    d0 = rng.uniform(1.0, 2.0)
    d1 = rng.uniform(0.5, 1.0)
    return Branch._trusted(np.array([L]), np.array([d0 / 2.0]), np.array([d1 / 2.0]), type_)


def _soma():
    return Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma")


def binary_tree_morphology(seed: int = 0, n_splits: int = 31) -> Morphology:
    rng = np.random.default_rng(seed)
    parents = [-1, 0]
    leaves = [1]
    for _ in range(n_splits):
        j = int(rng.integers(len(leaves)))
        leaf = leaves.pop(j)
        for _k in range(2):
            parents.append(leaf)
            leaves.append(len(parents) - 1)
    m = Morphology.from_root(_soma())
    for b in range(1, len(parents)):
        m.attach(parents[b], _dendrite(rng, "dendrite"), parent_x=1.0)
    return m


def grown_tree_morphology(rng, n_branches: int, p_split: float = 0.5) -> Morphology:
    parents = [-1, 0, 0]
    types = ["soma", "apical_dendrite", "basal_dendrite"]
    leaves = [1, 2]
    while len(parents) < n_branches:
        j = int(rng.integers(len(leaves)))
        leaf = leaves.pop(j)
        split = (n_branches - len(parents) >= 2) and (rng.random() < p_split)
        for _k in range(2 if split else 1):
            parents.append(leaf)
            types.append(types[leaf])
            leaves.append(len(parents) - 1)
    m = Morphology.from_root(_soma())
    for b in range(1, len(parents)):
        m.attach(parents[b], _dendrite(rng, types[b]), parent_x=1.0)
    return m


def single_tree_cell(seed: int = 0) -> CellModel:
    morph = binary_tree_morphology(seed)
    nb = morph.n_branches
    soma = np.asarray([b.type == "soma" for b in morph.branches])
    return CellModel(
        morph=morph, cv_per_branch=4, cm=1.0, ra=100.0, V0=-70.0,
        channels={
            "IL": {"g_branch": np.full(nb, 0.0003), "params": {"E": -70.0}},
            "Na_Ba2002": {"g_branch": np.where(soma, 0.12, 0.0), "params": {}},
            "KDR_Ba2002": {"g_branch": np.where(soma, 0.036, 0.0), "params": {}},
        },
        clamps=[(0, 0.5, 10.0, 50.0, 0.2)],
    ).validate()


def _ca_pool(mask_branch):
    return {"mask_branch": np.asarray(mask_branch, dtype=bool), "d": 1.0, "tau": 5.0, "C_rest": 2.4e-4,
            "Ci0": 2.4e-4, "Co": 2.0, "temp": celsius2kelvin(36.0)}


def uneven_tree_cell(index: int, cmin: int = 200, cmax: int = 2000) -> CellModel:
    rng = np.random.default_rng(index)
    c_target = int(rng.integers(cmin, cmax + 1))
    nbr = max(3, c_target // 4)
    morph = grown_tree_morphology(rng, nbr)
    nb = morph.n_branches
    dend = np.asarray([b.type != "soma" for b in morph.branches])
    z = rng.standard_normal(4)
    scale = np.exp(0.3 * z)
    amp = float(rng.uniform(0.1, 0.5))
    g = {k: BASE_DENSITIES[k] * s for k, s in zip(("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008"), scale)}
    return CellModel(
        morph=morph, cv_per_branch=4, cm=1.0, ra=100.0, V0=-70.0,
        channels={
            "Na_Ba2002": {"g_branch": np.full(nb, g["Na_Ba2002"]), "params": {}},
            "KDR_Ba2002": {"g_branch": np.full(nb, g["KDR_Ba2002"]), "params": {}},
            "IL": {"g_branch": np.full(nb, g["IL"]), "params": {"E": -70.0}},
            "CaL_IS2008": {"g_branch": np.where(dend, g["CaL_IS2008"], 0.0), "params": {}},
        },
        ca=_ca_pool(dend),
        clamps=[(0, 0.5, 10.0, CLAMP_FOREVER, amp)],
    ).validate()


def uneven_tree_batch(n_cells: int, start: int = 0, **kw) -> list:
    return [uneven_tree_cell(start + i, **kw) for i in range(n_cells)]


REGIONS = ("soma", "basal_dendrite", "apical_dendrite")


def region_tree_cell(index: int, densities=None, n_branches: int = 250, amp=None) -> CellModel:
    """``densities``: [3 regions, 4 channels] S/cm^2 in the order
    (Na_Ba2002, KDR_Ba2002, IL, CaL_IS2008); defaults to the base densities."""
    rng = np.random.default_rng(1_000_000 + index)
    morph = grown_tree_morphology(rng, n_branches)
    nb = morph.n_branches
    if densities is None:
        densities = np.tile(np.asarray([BASE_DENSITIES[k] for k in ("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")]),
                            (3, 1))
    densities = np.asarray(densities, dtype=np.float64).reshape(3, 4)
    reg = np.asarray([REGIONS.index(b.type) for b in morph.branches])
    if amp is None:
        amp = float(rng.uniform(0.1, 0.5))
    names = ("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")
    channels = {}
    for j, nm in enumerate(names):
        channels[nm] = {"g_branch": densities[reg, j], "params": {"E": -70.0} if nm == "IL" else {}}
    return CellModel(
        morph=morph, cv_per_branch=4, cm=1.0, ra=100.0, V0=-70.0, channels=channels,
        ca=_ca_pool(np.ones(nb, dtype=bool)), clamps=[(0, 0.5, 10.0, CLAMP_FOREVER, float(amp))],
    ).validate()


# ------------------------------------------------------- thalamocortical pop.
HTC_CHANNELS = {
    # README.md:52-71 HTC template (reference), densities in S/cm^2
    "Na_Ba2002": {"g_max": 0.09, "V_sh": -30.0},
    "K_Leak": {"g_max": 0.00001},
    "KDR_Ba2002": {"g_max": 0.01, "V_sh": -30.0, "q10": 2.0, "temp": celsius2kelvin(16.0)},
    "CaL_IS2008": {"g_max": 0.0005},
    "CaT_HM1992": {"g_max": 0.0021},
    "AHP_De1994": {"g_max": 0.0003},
    "HCN_HM1992": {"g_max": 0.00001, "E": -43.0},
    "IL": {"g_max": 0.0000075, "E": -70.0},
}
HTC_IONS = {"E_Na": 50.0, "E_K": -90.0, "ca": {"C_rest": 5e-5, "tau": 10.0, "d": 0.5, "Ci0": 2.4e-4, "Co": 2.0,
                                                "temp": celsius2kelvin(36.0)}}


def htc_population(n: int = 100_000, seed: int = 1):
    """Per-neuron lognormal (sigma 0.3) densities around the HTC template,
    V0 ~ U(-70, -60) mV and DC ~ U(0, 0.3) nA.  Returns a dict of arrays."""
    rng = np.random.default_rng(seed)
    names = list(HTC_CHANNELS)
    z = rng.standard_normal((len(names), n))
    g = {nm: HTC_CHANNELS[nm]["g_max"] * np.exp(0.3 * z[i]) for i, nm in enumerate(names)}
    V0 = rng.uniform(-70.0, -60.0, n)
    dc = rng.uniform(0.0, 0.3, n)
    return {"g": g, "V0": V0, "I_dc_nA": dc, "names": names}


_N144_MORPH = None


def n144_amplitudes(n: int):
    """Clamp amplitudes in nA: 0.7 for one cell, else 0.69 ... 0.71 evenly
    (``simulator_compare/common.py:58-64``)."""
    if n <= 0:
        raise ValueError("n must be positive")
    if n == 1:
        return [0.7]
    return [0.69 + 0.02 * i / (n - 1) for i in range(n)]


def n144_hh_cell(amp: float = 0.7) -> CellModel:
    global _N144_MORPH
    from .swc import read_swc

    if _N144_MORPH is None:
        _N144_MORPH = read_swc(N144_SWC, mode="neuron")
    morph = _N144_MORPH
    nb = morph.n_branches
    t = celsius2kelvin(6.3)
    hh = {"temp": t, "temp_ref": t}
    return CellModel(
        morph=morph, cv_per_branch=4, cm=1.0, ra=1000.0, V0=-62.0, E_Na=50.0, E_K=-77.0,
        channels={
            "Na_HH1952": {"g_branch": np.full(nb, 0.12), "params": dict(hh)},
            "K_HH1952": {"g_branch": np.full(nb, 0.036), "params": dict(hh)},
            "IL": {"g_branch": np.full(nb, 0.0003), "params": {"E": -54.3}},
        },
        clamps=[(0, 0.85, 3.0, 2.0, float(amp))],
    ).validate()
