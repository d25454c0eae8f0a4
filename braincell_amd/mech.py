"""Mechanism declarations painted or placed on a :class:`~braincell_amd.cell.Cell`.

Mirrors ``braincell.mech``: ``Channel`` / ``Ion`` (``mech/_density.py:315-496``),
``CableProperty`` (``mech/_cable.py:42``) and ``CurrentClamp``
(``mech/_point.py:86-141``).  Values are plain floats in fixed units:

- conductance densities ``g_max`` in S/cm^2 (as the reference README paints),
- voltages in mV, times in ms, currents in nA,
- capacitance in uF/cm^2, axial resistivity in ohm cm,
- concentrations in mM, lengths in um, temperatures in K.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from .channels import CHANNELS, IONS, celsius2kelvin

__all__ = ["Channel", "Ion", "CableProperty", "CurrentClamp", "StateProbe", "MechanismProbe", "CurrentProbe"]


@dataclass(frozen=True)
class CableProperty:
    """Passive cable constants of a region (``mech/_cable.py:42-120``)."""

    resting_potential: float = -65.0
    membrane_capacitance: float = 1.0
    axial_resistivity: float = 100.0
    temperature: float = celsius2kelvin(36.0)

    def __post_init__(self):
        if self.membrane_capacitance <= 0 or self.axial_resistivity <= 0:
            raise ValueError("membrane_capacitance and axial_resistivity must be > 0")


@dataclass(frozen=True)
class Channel:
    """Distributed ion-channel declaration: a catalogue name plus parameters.

    ``Channel("Na_Ba2002", g_max=0.12)``; the ``I`` prefix of the reference's
    registry aliases (``"INa_Ba2002"``) is accepted.
    """

    class_name: str
    params: dict = field(default_factory=dict)

    def __init__(self, class_name: str, **params):
        name = class_name
        if name not in CHANNELS and name.startswith("I") and name[1:] in CHANNELS:
            name = name[1:]
        if name not in CHANNELS:
            raise ValueError(f"Unknown channel {class_name!r}. Available: {sorted(CHANNELS)}")
        unknown = set(params) - set(CHANNELS[name].params)
        if unknown:
            raise TypeError(f"{name}: unknown parameter(s) {sorted(unknown)}")
        object.__setattr__(self, "class_name", name)
        object.__setattr__(self, "params", dict(params))

    def resolved(self) -> dict:
        out = dict(CHANNELS[self.class_name].params)
        out.update(self.params)
        return out


@dataclass(frozen=True)
class Ion:
    """Distributed ion-species declaration (``SodiumFixed``, ``PotassiumFixed``,
    ``CalciumFixed``, ``CalciumDetailed``)."""

    class_name: str
    params: dict = field(default_factory=dict)

    def __init__(self, class_name: str, **params):
        if class_name not in IONS:
            raise ValueError(f"Unknown ion {class_name!r}. Available: {sorted(IONS)}")
        unknown = set(params) - set(IONS[class_name])
        if unknown:
            raise TypeError(f"{class_name}: unknown parameter(s) {sorted(unknown)}")
        object.__setattr__(self, "class_name", class_name)
        object.__setattr__(self, "params", dict(params))

    def resolved(self) -> dict:
        out = dict(IONS[self.class_name])
        out.update(self.params)
        return out


@dataclass(frozen=True)
class CurrentClamp:
    """Piecewise-constant current clamp (``mech/_point.py:86-141``).

    Segment ``k`` is active while ``delay + sum(durations[:k]) <= t <
    delay + sum(durations[:k+1])`` (``_compute/layouts.py:960-982``).
    """

    delay: float = 0.0
    durations: tuple = (1.0,)
    amplitudes: tuple = (0.0,)

    def __init__(self, delay=0.0, durations=1.0, amplitudes=0.0):
        durs = tuple(float(d) for d in (durations if hasattr(durations, "__len__") else (durations,)))
        amps = tuple(float(a) for a in (amplitudes if hasattr(amplitudes, "__len__") else (amplitudes,)))
        if len(durs) != len(amps):
            raise ValueError("durations and amplitudes must have the same length")
        if any(d <= 0 for d in durs):
            raise ValueError("durations must be > 0")
        object.__setattr__(self, "delay", float(delay))
        object.__setattr__(self, "durations", durs)
        object.__setattr__(self, "amplitudes", amps)

    def segments(self):
        """(start, end, amplitude) in absolute ms."""
        out = []
        s = self.delay
        for d, a in zip(self.durations, self.amplitudes):
            out.append((s, s + d, a))
            s += d
        return out


@dataclass(frozen=True)
class StateProbe:
    """Records the membrane voltage at its locations after every step
    (``mech/_point.py:314``, ``_multi_compartment/probes.py:124-134``)."""

    name: str
    field: str = "v"

    def __post_init__(self):
        if self.field != "v":
            raise ValueError(f"Unsupported StateProbe field {self.field!r}.")


@dataclass(frozen=True)
class MechanismProbe:
    """Records a state of a painted mechanism after every step
    (``probes.py:137-180``): ``mechanism`` is a channel class name with
    ``field`` one of its gates (``"p"``, ``"q"``), or ``"CalciumDetailed"``
    with ``field="C"`` (the pool concentration, mM)."""

    name: str
    mechanism: str
    field: str


@dataclass(frozen=True)
class CurrentProbe:
    """Records a membrane current density (uA/cm^2, outward negative as in
    the reference's ``g (E - V)``) after every step (``probes.py:183-236``):
    one channel's (``mechanism``) or an ion's total (``ion="Ca"``)."""

    name: str
    mechanism: str = None
    ion: str = None

    def __post_init__(self):
        if (self.mechanism is None) == (self.ion is None):
            raise ValueError(f"Probe {self.name!r} must define exactly one of 'mechanism' and 'ion'.")
