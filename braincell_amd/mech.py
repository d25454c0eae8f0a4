"""Mechanism declarations painted or placed on a :class:`~braincell_amd.cell.Cell`.

Mirrors ``braincell.mech``: ``Channel`` / ``Ion`` (``mech/_density.py:315-496``),
``CableProperty`` (``mech/_cable.py:42``), ``CurrentClamp``
(``mech/_point.py:86-141``) and ``Synapse`` (``mech/_point.py:389-440``).  Values are plain floats in fixed units:

- conductance densities ``g_max`` in S/cm^2 (as the reference README paints),
  synaptic conductances and weights in uS,
- voltages in mV, times in ms, currents in nA,
- capacitance in uF/cm^2, axial resistivity in ohm cm,
- concentrations in mM, lengths in um, temperatures in K.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .channels import CHANNELS, IONS, celsius2kelvin

__all__ = ["Channel", "Ion", "CableProperty", "CurrentClamp", "NetStim", "StateProbe", "MechanismProbe",
           "CurrentProbe", "Synapse"]


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
class NetStim:
    """Deterministic spike source at a point (``mech/_point.py:252-311``,
    NEURON's NetStim with noise 0): events at start + k * interval ms for
    k = 0 .. number - 1, each a drive of ``weight`` (the payload units of the
    synapses at the same point: uS for ExpSyn / Exp2Syn, dimensionless for
    AMPA / GABAa / NMDA) for every synapse placed at its CV.

    ``fired(t)`` is the reference's test (``_compute/layouts.py:906-941``):
    with local = t - start and k = round(local / interval) (half to even),
    an event is due when local >= 0, |local - k interval| <= 1e-9 and
    0 <= k < number; evaluated in float64 here."""

    start: float
    number: int
    interval: float
    noise: float = 0.0
    weight: float = 1.0
    name: str = None

    def __post_init__(self):
        for what in ("start", "interval"):
            v = getattr(self, what)
            if isinstance(v, bool) or not isinstance(v, (int, float, np.integer, np.floating)):
                raise TypeError(f"NetStim.{what} must be a scalar time in ms, got {v!r}.")
            object.__setattr__(self, what, float(v))
        if self.interval <= 0.0:
            raise ValueError(f"NetStim.interval must be > 0, got {self.interval!r}.")
        if isinstance(self.number, bool) or not isinstance(self.number, int):
            raise TypeError(f"NetStim.number must be int, got {self.number!r}.")
        if self.number < 0:
            raise ValueError(f"NetStim.number must be >= 0, got {self.number!r}.")
        if isinstance(self.noise, bool) or not isinstance(self.noise, (int, float)):
            raise TypeError(f"NetStim.noise must be a real number, got {self.noise!r}.")
        if float(self.noise) != 0.0:
            raise ValueError(f"NetStim.noise={self.noise!r} is not supported yet; use noise=0.0.")
        if isinstance(self.weight, bool) or not isinstance(self.weight, (int, float)):
            raise TypeError(f"NetStim.weight must be a real number, got {self.weight!r}.")
        if self.name is not None and (not isinstance(self.name, str) or not self.name):
            raise ValueError(f"NetStim.name must be a non-empty string or None, got {self.name!r}.")

    @property
    def instance_name(self) -> str:
        return self.name if self.name is not None else "NetStim"

    def fired(self, t_ms) -> np.ndarray:
        local = np.asarray(t_ms, dtype=np.float64) - self.start
        k = np.rint(local / self.interval)
        on_grid = np.abs(local - k * self.interval) <= 1e-9
        return (local >= 0.0) & on_grid & (k >= 0) & (k < self.number)


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


_SYNAPSE_DEFAULTS = {
    "ExpSyn": {"tau": 0.1, "e": 0.0, "weight": 1.0},
    "Exp2Syn": {"tau1": 0.1, "tau2": 10.0, "e": 0.0, "weight": 1.0},
    # synapse/markov.py:227-236, 271-280, 315-326 (rates 1/ms, g_max mS/cm^2, E_rev mV)
    "AMPA": {"alpha": 0.98, "beta": 0.18, "T": 0.5, "g_max": 1.0, "E_rev": 0.0},
    "GABAa": {"alpha": 0.53, "beta": 0.18, "T": 1.0, "g_max": 1.0, "E_rev": -70.0},
    "NMDA": {"alpha1": 2.0, "beta1": 0.01, "alpha2": 1.0, "beta2": 0.5, "T": 1.0, "g_max": 1.0, "E_rev": 0.0},
}
# kinds of the kernel's synapse update (bc_synapse_desc.par[2])
SYN_EXP, SYN_MARKOV1, SYN_MARKOV2 = 0, 1, 2


class Synapse:
    """Registry-keyed synapse placed with ``Cell.place`` (``mech/_point.py:
    389-440``): ``Synapse("ExpSyn", tau=2.0, e=0.0, weight=1.0, name="exp")``.

    Types (``synapse/markov.py:65-360``):

    - ``ExpSyn`` (NEURON's expsyn): g' = -g / tau, an event adds its weight
      to g; current g (V - e) in nA (outward positive);
    - ``Exp2Syn`` (exp2syn): A' = -A / tau1, B' = -B / tau2, g = B - A, an
      event adds weight * factor to A and B (factor normalises the peak of g
      to the weight; tau1 is kept in [1e-9, 0.9999] * tau2);
    - ``AMPA`` / ``GABAa``: g' = alpha D T (1 - g) - beta g, with D the
      summed weighted events of the step (a transmitter pulse one step
      long), current density g_max g (E_rev - V) in mS/cm^2 * mV (inward
      positive);
    - ``NMDA``: g' = alpha1 x (1 - g) - beta1 g, x' = alpha2 D T (1 - x) -
      beta2 x, same current density.

    The Markov types integrate with ind_exp_euler after the voltage step
    (every state from the step-start values) and have no default weight:
    their connections give one (dimensionless).  A placement over several
    locations is one pool; a contact's ``synapse_index`` picks the location.
    """

    __slots__ = ("synapse_type", "params", "name")

    def __init__(self, synapse_type: str, /, *, name: str = None, **params):
        if synapse_type not in _SYNAPSE_DEFAULTS:
            raise ValueError(f"Unknown synapse type {synapse_type!r}; supported: {sorted(_SYNAPSE_DEFAULTS)}.")
        unknown = set(params) - set(_SYNAPSE_DEFAULTS[synapse_type])
        if unknown:
            raise ValueError(f"{synapse_type} has no parameters {sorted(unknown)}.")
        self.synapse_type = synapse_type
        self.params = {k: float(v) for k, v in params.items()}
        self.name = name

    @property
    def instance_name(self) -> str:
        return self.name if self.name is not None else self.synapse_type

    def resolved(self) -> dict:
        return {**_SYNAPSE_DEFAULTS[self.synapse_type], **self.params}

    def effective_tau1(self) -> float:
        """``Exp2Syn._effective_tau1`` (``synapse/markov.py:152-156``)."""
        p = self.resolved()
        ratio = p["tau1"] / p["tau2"]
        if ratio > 0.9999:
            return p["tau2"] * 0.9999
        if ratio < 1e-9:
            return p["tau2"] * 1e-9
        return p["tau1"]

    def factor(self) -> float:
        """``Exp2Syn._compute_factor`` (``synapse/markov.py:158-162``)."""
        import math

        p = self.resolved()
        t1, t2 = self.effective_tau1(), p["tau2"]
        tp = (t1 * t2) / (t2 - t1) * math.log(t2 / t1)
        return 1.0 / (-math.exp(-(tp / t1)) + math.exp(-(tp / t2)))

    @property
    def markov(self) -> bool:
        return self.synapse_type in ("AMPA", "GABAa", "NMDA")

    def kernel_par(self, dt: float, cm: float, area_cm2: float):
        """The kernel's 8 parameters of one synapse at a CV of membrane
        capacitance ``cm`` (uF/cm^2) and area ``area_cm2``: k (dt/Cm times
        the state -> mS/cm^2 conductance factor), reversal potential, kind
        (SYN_EXP, SYN_MARKOV1, SYN_MARKOV2) and five kind-specific terms:
        exponential (fa, fb, ka, kb, 0); one-state Markov (alpha T, beta,
        0, 0, 0); NMDA (alpha1, beta1, alpha2 T, beta2, 0)."""
        p = self.resolved()
        if not self.markov:
            e, fa, fb, ka, kb = self.kernel_terms(dt)
            return (dt / cm * 1e-3 / area_cm2, e, float(SYN_EXP), fa, fb, ka, kb, 0.0)
        k = dt / cm * p["g_max"]
        if self.synapse_type == "NMDA":
            return (k, p["E_rev"], float(SYN_MARKOV2), p["alpha1"], p["beta1"], p["alpha2"] * p["T"], p["beta2"], 0.0)
        return (k, p["E_rev"], float(SYN_MARKOV1), p["alpha"] * p["T"], p["beta"], 0.0, 0.0, 0.0)

    def kernel_terms(self, dt: float):
        """(e, fa, fb, ka, kb): event gains of the states A and B and their
        exponential-Euler decay fractions over ``dt`` (y -= k y with k =
        dt exprel(-dt / tau) / tau, ``quad/_exp_euler.py:462-465``)."""
        import math

        def k(tau):
            x = -dt / tau
            ex = 1.0 + x * (0.5 + x / 6.0) if abs(x) < 1e-8 else math.expm1(x) / x
            return dt * ex / tau

        p = self.resolved()
        if self.markov:
            raise ValueError(f"{self.synapse_type} is not an exponential synapse")
        if self.synapse_type == "ExpSyn":
            return p["e"], 0.0, 1.0, 0.0, k(p["tau"])
        f = self.factor()
        return p["e"], f, f, k(self.effective_tau1()), k(p["tau2"])

    def __repr__(self):
        return f"Synapse({self.synapse_type!r}, name={self.name!r}, **{self.params!r})"
