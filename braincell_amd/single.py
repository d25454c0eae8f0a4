"""``SingleCompartment`` populations on the point-neuron HIP kernel.

Mirrors ``braincell.SingleCompartment`` (``_single_compartment/base.py:33-283``):
a population of ``size`` point neurons with length/radius (area
``2 pi r L``, ``base.py:150-152``), capacitance ``C``, threshold ``V_th`` and a
solver resolved through the integrator registry; ``update(I_ext)`` runs the
solver for one step.  The kernel implements the thalamocortical channel set
(``README.md:52-71`` of the reference): Na_Ba2002, K_Leak, KDR_Ba2002,
CaL_IS2008, CaT_HM1992, AHP_De1994, HCN_HM1992, IL and a CalciumDetailed
pool; channels left out get zero conductance.  Solvers: ``ind_exp_euler_mi355x``
(default), ``backward_euler_mi355x`` and ``implicit_euler_mi355x``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import rates
from ._native import chan_const_type, check, lib
from .channels import CHANNELS, FARADAY, GAS_CONSTANT, IONS, celsius2kelvin
from .quad import get_integrator

__all__ = ["SingleCompartment", "PointDesc", "POINT_CHANNELS", "IND_EXP_EULER", "BACKWARD_EULER", "IMPLICIT_EULER"]

POINT_CHANNELS = ("Na_Ba2002", "K_Leak", "KDR_Ba2002", "CaL_IS2008", "CaT_HM1992", "AHP_De1994", "HCN_HM1992", "IL")


class PointDesc(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("n", ctypes.c_int64), ("n_steps", ctypes.c_int32),
                ("dt", ctypes.c_double), ("Cm", ctypes.c_double), ("V_th", ctypes.c_double),
                ("chan_const", ctypes.c_void_p), ("V", ctypes.c_void_p), ("state", ctypes.c_void_p),
                ("g", ctypes.c_void_p), ("I_ext", ctypes.c_void_p), ("trace", ctypes.c_void_p),
                ("n_trace", ctypes.c_int64), ("spikes", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("method", ctypes.c_int32)]

# bc_point_desc.method values
IND_EXP_EULER, BACKWARD_EULER, IMPLICIT_EULER = 0, 1, 2


def _phi(q10, t, tr):
    return q10 ** ((t - tr) / 10.0)


class SingleCompartment:
    """Population of point neurons integrated by a registered solver.

    ``channels``: name -> dict with ``g_max`` (S/cm^2, scalar or [size]) and
    optional non-density parameters; ``E_Na`` / ``E_K`` fixed reversal
    potentials (mV); ``ca``: CalciumDetailed parameters.
    """

    def __init__(self, size: int, *, channels: dict, E_Na: float = 50.0, E_K: float = -90.0, ca: dict = None,
                 length: float = 10.0, radius: float = 5.0, C: float = 1.0, V_th: float = 0.0, V0=-65.0,
                 solver="ind_exp_euler_mi355x", precision: int = 32, device="cuda"):
        self.size = int(size)
        unknown = set(channels) - set(POINT_CHANNELS)
        if unknown:
            raise ValueError(f"channels {sorted(unknown)} are not in the point-neuron kernel")
        self.channels = channels
        self.E_Na, self.E_K = float(E_Na), float(E_K)
        self.ca = dict(IONS["CalciumDetailed"], **(ca or {}))
        self.length, self.radius, self.C, self.V_th = float(length), float(radius), float(C), float(V_th)
        self.V0 = V0
        self.solver = get_integrator(solver)
        # the kernel method ``run`` uses; the registered point solvers carry it
        self.method = getattr(self.solver, "point_method", IND_EXP_EULER)
        self.precision = int(precision)
        self.device = device
        self._dev = None

    @property
    def area_cm2(self) -> float:
        return 2.0 * np.pi * self.radius * self.length * 1e-8

    def _params(self, name):
        p = dict(CHANNELS[name].params)
        p.update({k: v for k, v in self.channels.get(name, {}).items() if k != "g_max"})
        return p

    def chan_const(self):
        na, kd = self._params("Na_Ba2002"), self._params("KDR_Ba2002")
        cl, ct = self._params("CaL_IS2008"), self._params("CaT_HM1992")
        ah, hc, il = self._params("AHP_De1994"), self._params("HCN_HM1992"), self._params("IL")
        ca = self.ca
        return {
            "E_Na": self.E_Na, "E_K": self.E_K, "E_IL": il["E"], "na_vsh": na["V_sh"],
            "na_phi": _phi(na["q10"], na["temp"], na["temp_ref"]), "kdr_vsh": kd["V_sh"],
            "kdr_phi": _phi(kd["q10"], kd["temp"], kd["temp_ref"]), "cal_vsh": cl["V_sh"],
            "cal_phi_p": _phi(cl["q10_p"], cl["temp"], cl["temp_ref_p"]),
            "cal_phi_q": _phi(cl["q10_q"], cl["temp"], cl["temp_ref_q"]), "cat_vsh": ct["V_sh"],
            "cat_phi_p": _phi(ct["q10_p"], ct["temp"], ct["temp_ref_p"]),
            "cat_phi_q": _phi(ct["q10_q"], ct["temp"], ct["temp_ref_q"]),
            "ahp_alpha": ah["alpha"], "ahp_beta": ah["beta"], "ahp_n": ah["n"], "ahp_phi": ah["phi"],
            "hcn_E": hc["E"], "hcn_phi": _phi(hc["q10"], hc["temp"], hc["temp_ref"]),
            "ca_drive": 10.0 / (2.0 * FARADAY * float(ca["d"])), "ca_tau_inv": 1.0 / float(ca["tau"]),
            "ca_rest": float(ca["C_rest"]), "ca_Co": float(ca["Co"]),
            "ca_rtzf": 1e3 * GAS_CONSTANT * float(ca.get("temp", celsius2kelvin(36.0))) / (2.0 * FARADAY),
            "E_Ca_fixed": IONS["CalciumFixed"]["E"],
        }

    def initial_state(self, V0=None):
        """reset_state: V = V0, gates at steady state, Ci at its initializer."""
        V = np.broadcast_to(np.asarray(self.V0 if V0 is None else V0, dtype=np.float64), (self.size,)).copy()
        cc = self.chan_const()
        ci = np.full(self.size, float(self.ca["Ci0"]))
        S = np.empty((10, self.size))
        S[0], S[1] = rates.na_steady(V, cc["na_vsh"])
        S[2] = rates.kdr_steady(V, cc["kdr_vsh"])
        S[3], S[4] = rates.cal_steady(V, cc["cal_vsh"])
        S[5], S[6] = rates.cat_steady(V, cc["cat_vsh"])
        S[7] = rates.ahp_steady(ci, cc["ahp_alpha"], cc["ahp_beta"], cc["ahp_n"])
        S[8] = rates.hcn_steady(V)
        S[9] = ci
        return V, S

    def g_table(self):
        g = np.zeros((8, self.size))
        for k, name in enumerate(POINT_CHANNELS):
            if name in self.channels:
                g[k] = np.broadcast_to(np.asarray(self.channels[name]["g_max"], dtype=np.float64), (self.size,)) * 1e3
        return g

    # ------------------------------------------------------------- device
    def init_state(self):
        import torch

        lib()
        dt = torch.float32 if self.precision == 32 else torch.float64
        dev = torch.device(self.device)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=dt)
        V, S = self.initial_state()
        self._dev = {
            "V": t(V), "S": t(S.reshape(-1)), "g": t(self.g_table().reshape(-1)),
            "I": torch.zeros(self.size, dtype=dt, device=dev),
            "spikes": torch.zeros(self.size, dtype=torch.int32, device=dev),
            "status": torch.zeros(1, dtype=torch.int32, device=dev),
        }
        self._dtype = dt
        self._cc = chan_const_type(self.precision)(**{k: float(v) for k, v in self.chan_const().items()})
        return self

    def reset_state(self, V0=None):
        if self._dev is None:
            self.init_state()
        import torch

        V, S = self.initial_state(V0)
        self._dev["V"].copy_(torch.as_tensor(V))
        self._dev["S"].copy_(torch.as_tensor(S.reshape(-1)))
        self._dev["spikes"].zero_()
        self._dev["status"].zero_()

    def set_input(self, I_nA):
        """Injected current (nA), converted to a density through the area."""
        import torch

        if self._dev is None:
            self.init_state()
        dens = np.broadcast_to(np.asarray(I_nA, dtype=np.float64), (self.size,)) * 1e-3 / self.area_cm2
        self._dev["I"].copy_(torch.as_tensor(dens))

    @property
    def V(self):
        return self._dev["V"]

    @property
    def spikes(self):
        return self._dev["spikes"]

    def desc(self, dt, n_steps, trace=None, n_trace=0, method=None):
        d = PointDesc()
        d.method = self.method if method is None else int(method)
        dv = self._dev
        p = lambda x: ctypes.c_void_p(x.data_ptr())
        d.precision, d.n, d.n_steps, d.dt, d.Cm, d.V_th = self.precision, self.size, int(n_steps), float(dt), self.C, self.V_th
        d.chan_const = ctypes.cast(ctypes.pointer(self._cc), ctypes.c_void_p)
        d.V, d.state, d.g, d.I_ext = p(dv["V"]), p(dv["S"]), p(dv["g"]), p(dv["I"])
        d.trace = p(trace) if trace is not None else None
        d.n_trace = int(n_trace) if trace is not None else 0
        d.spikes, d.status = p(dv["spikes"]), p(dv["status"])
        return d

    def advance(self, t, dt, n_steps=1, I_ext=None, trace=None, n_trace=0, method=None):
        """``n_steps`` steps of ``dt`` ms with kernel ``method`` (``IND_EXP_EULER``,
        ``BACKWARD_EULER`` or ``IMPLICIT_EULER``; default: the solver's)."""
        import torch

        if self._dev is None:
            self.init_state()
        if I_ext is not None:
            self.set_input(I_ext)
        if trace is not None and (trace.dtype != self._dtype or trace.numel() < n_steps * n_trace):
            raise ValueError("trace must hold n_steps * n_trace values of the population dtype")
        d = self.desc(dt, n_steps, trace, n_trace, method)
        check(lib().bc_point_run(ctypes.byref(d), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return trace

    def update(self, I_ext=None):
        """One step through the configured solver (``base.py:252-278``)."""
        return self.solver(self, I_ext)

    def run(self, *, dt: float, duration: float, n_trace: int = 0, chunk: int = 4000):
        import torch

        if self._dev is None:
            self.init_state()
        n = int(np.arange(0.0, float(duration), float(dt)).shape[0])
        trace = torch.empty((n, n_trace), dtype=self._dtype, device=self._dev["V"].device) if n_trace else None
        done = 0
        while done < n:
            k = min(chunk, n - done)
            self.advance(done * dt, dt, k, trace=trace[done:] if trace is not None else None, n_trace=n_trace)
            done += k
        torch.cuda.synchronize()
        if int(self._dev["status"].item()):
            from ._native import NativeError

            raise NativeError(4, "non-finite membrane voltage detected")
        return trace
