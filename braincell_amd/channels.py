"""Catalogue of the channels and ions the HIP kernels implement.

Each entry records the reference class it mirrors, its parameter defaults
(same defaults as the reference constructors, with ``g_max`` converted to
S/cm^2), its gate structure and the kernel's channel bit.  The rate
functions themselves are compiled into ``csrc/bc_channels.h``.
"""

from __future__ import annotations

from dataclasses import dataclass

__all__ = ["CHANNELS", "IONS", "ChannelInfo", "celsius2kelvin", "FARADAY", "GAS_CONSTANT", "channel_mask"]

FARADAY = 96485.33212  # C/mol, CODATA 2018
GAS_CONSTANT = 8.314462618  # J/(mol K), CODATA 2018


def celsius2kelvin(c: float) -> float:
    return float(c) + 273.15


@dataclass(frozen=True)
class ChannelInfo:
    bit: int
    ion: str  # "Na", "K", "Ca", "KCa", "none"
    gates: tuple  # (name, power, form) with form "ab" or "it"
    params: dict
    reference: str


CHANNELS = {
    "Na_Ba2002": ChannelInfo(
        1 << 0, "Na", (("p", 3, "ab"), ("q", 1, "ab")),
        {"g_max": 0.09, "temp": celsius2kelvin(36.0), "q10": 3.0, "temp_ref": celsius2kelvin(36.0), "V_sh": -50.0},
        "braincell/channel/sodium.py:58-170"),
    "KDR_Ba2002": ChannelInfo(
        1 << 1, "K", (("p", 4, "ab"),),
        {"g_max": 0.01, "temp": celsius2kelvin(36.0), "q10": 3.0, "temp_ref": celsius2kelvin(36.0), "V_sh": -50.0},
        "braincell/channel/potassium.py:95-214"),
    "IL": ChannelInfo(
        1 << 2, "none", (), {"g_max": 0.0001, "E": -70.0}, "braincell/channel/leaky.py:144-200"),
    "K_Leak": ChannelInfo(
        1 << 3, "K", (), {"g_max": 0.000005}, "braincell/channel/potassium.py:1167-1235"),
    "CaL_IS2008": ChannelInfo(
        1 << 4, "Ca", (("p", 2, "it"), ("q", 1, "it")),
        {"g_max": 0.001, "temp": celsius2kelvin(36.0), "q10_p": 3.55, "temp_ref_p": celsius2kelvin(24.0),
         "q10_q": 3.0, "temp_ref_q": celsius2kelvin(24.0), "V_sh": 0.0},
        "braincell/channel/calcium.py:790-920"),
    "CaT_HM1992": ChannelInfo(
        1 << 5, "Ca", (("p", 2, "it"), ("q", 1, "it")),
        {"g_max": 0.002, "temp": celsius2kelvin(36.0), "q10_p": 3.55, "temp_ref_p": celsius2kelvin(24.0),
         "q10_q": 3.0, "temp_ref_q": celsius2kelvin(24.0), "V_sh": -3.0},
        "braincell/channel/calcium.py:222-373"),
    "AHP_De1994": ChannelInfo(
        1 << 6, "KCa", (("p", 2, "ab"),),
        {"g_max": 0.01, "n": 2.0, "alpha": 48.0, "beta": 0.09, "phi": 1.0},
        "braincell/channel/potassium_calcium.py:78-185"),
    "Na_HH1952": ChannelInfo(
        1 << 9, "Na", (("p", 3, "ab"), ("q", 1, "ab")),
        {"g_max": 0.12, "temp": celsius2kelvin(36.0), "q10": 3.0, "temp_ref": celsius2kelvin(36.0), "V_sh": -45.0},
        "braincell/channel/sodium.py:293-398"),
    "K_HH1952": ChannelInfo(
        1 << 10, "K", (("p", 4, "ab"),),
        {"g_max": 0.01, "temp": celsius2kelvin(36.0), "q10": 3.0, "temp_ref": celsius2kelvin(36.0), "V_sh": -45.0},
        "braincell/channel/potassium.py:341-456"),
    "HCN_HM1992": ChannelInfo(
        1 << 7, "none", (("p", 1, "it"),),
        {"g_max": 0.01, "E": 43.0, "temp": celsius2kelvin(36.0), "q10": 1.0, "temp_ref": celsius2kelvin(36.0)},
        "braincell/channel/hyperpolarization_activated.py:45-132"),
}

# Bit used by the kernels for the CalciumDetailed pool (not a channel).
CA_POOL_BIT = 1 << 8

IONS = {
    "SodiumFixed": {"E": 50.0},
    "PotassiumFixed": {"E": -95.0},
    "CalciumFixed": {"E": 120.0},
    "CalciumDetailed": {"temp": celsius2kelvin(36.0), "d": 1.0, "tau": 5.0, "C_rest": 2.4e-4,
                        "Co": 2.0, "Ci0": 2.4e-4},
}


def channel_mask(names, ca_pool: bool) -> int:
    m = 0
    for n in names:
        m |= CHANNELS[n].bit
    if ca_pool:
        m |= CA_POOL_BIT
    return m
