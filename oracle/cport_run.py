"""ctypes driver of the FP64 C port (test infrastructure / CPU baseline).

Cells must use the channel set Na_Ba2002 + KDR_Ba2002 + IL + CaL_IS2008 with
a CalciumDetailed pool (the batched uneven-tree workload)."""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import channels as ch
from .staggered import StaggeredOracle

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cport", "libbcport.so")
ORDER = ("Na_Ba2002", "KDR_Ba2002", "IL", "K_Leak", "CaL_IS2008", "CaT_HM1992", "AHP_De1994", "HCN_HM1992")

I = ctypes.POINTER(ctypes.c_int)
D = ctypes.POINTER(ctypes.c_double)


class PortCell(ctypes.Structure):
    _fields_ = [("C", ctypes.c_int), ("P", ctypes.c_int), ("n_edges", ctypes.c_int), ("n_levels", ctypes.c_int),
                ("n_jump", ctypes.c_int), ("dyn", I), ("parent", I), ("diag", D), ("lowers", D), ("uppers", D),
                ("edges", I), ("level_off", I), ("jumps", I), ("cm", D), ("g", D), ("camask", D), ("clamp", D),
                ("clamp_on", ctypes.c_int), ("clamp_off", ctypes.c_int), ("soma_cv", ctypes.c_int), ("V", D),
                ("S", D)]


class PortConst(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "E_Na", "E_K", "E_IL", "na_vsh", "na_phi", "kdr_vsh", "kdr_phi", "cal_vsh", "cal_phi_p", "cal_phi_q",
        "ca_drive", "ca_tau_inv", "ca_rest", "ca_Co", "ca_rtzf")]


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: run python oracle/build_cport.py")
    L = ctypes.CDLL(LIB)
    L.port_run.argtypes = [ctypes.c_int, ctypes.POINTER(PortCell), ctypes.POINTER(PortConst), ctypes.c_double,
                           ctypes.c_int, ctypes.c_int, D, ctypes.c_int, ctypes.c_int]
    L.port_run.restype = ctypes.c_int
    L.port_max_threads.restype = ctypes.c_int
    return L


def _first_step_at(t, dt):
    if t <= 0:
        return 0
    if t / dt > 2 ** 31 - 1:
        return 2 ** 31 - 1
    n = max(0, int(np.ceil(t / dt)) - 1)
    while n * dt < t:
        n += 1
    while n > 0 and (n - 1) * dt >= t:
        n -= 1
    return n


class PortBatch:
    """Builds the per-cell inputs from the oracle's discretisation."""

    def __init__(self, models, dt, soma_cvs):
        self.keep = []
        self.cells = (PortCell * len(models))()
        self.dt = float(dt)
        for n, (m, scv) in enumerate(zip(models, soma_cvs)):
            o = StaggeredOracle(m)
            o.reset_state(m.V0)
            st = o.st
            a = lambda x, t: np.ascontiguousarray(np.asarray(x, dtype=t))
            dyn, par = a(st.dynamic_rows, np.int32), a(st.parent_rows, np.int32)
            diag, lo, up = a(st.diag, np.float64), a(st.lowers, np.float64), a(st.uppers, np.float64)
            edges, lvl, jumps = a(st.edges.reshape(-1), np.int32), a(st.level_offsets, np.int32), a(st.backsub, np.int32)
            g = np.zeros((o.C, 8))
            for j, nm in enumerate(ORDER):
                if nm in o.g:
                    g[:, j] = o.g[nm]
            g = a(g.reshape(-1), np.float64)
            camask = a(o.ca_mask.astype(np.float64), np.float64)
            clamp = np.zeros(o.C)
            on = off = 0
            for cv, t0, dur, amp in o.mid_clamps:
                clamp[cv] = amp * 1e-3 / o.area_cm2[cv]
                on, off = _first_step_at(t0, dt), _first_step_at(t0 + dur, dt)
            clamp = a(clamp, np.float64)
            V = a(o.V, np.float64)
            S = a(np.concatenate([o.gates["Na_Ba2002"][0], o.gates["Na_Ba2002"][1], o.gates["KDR_Ba2002"][0],
                                  o.gates["CaL_IS2008"][0], o.gates["CaL_IS2008"][1], o.Ci]), np.float64)
            arrs = (dyn, par, diag, lo, up, edges, lvl, jumps, a(o.cm, np.float64), g, camask, clamp, V, S)
            self.keep.append(arrs)
            ip = lambda x: x.ctypes.data_as(I)
            dp = lambda x: x.ctypes.data_as(D)
            self.cells[n] = PortCell(o.C, st.n_point, st.edges.shape[0], len(st.level_offsets) - 1,
                                     st.backsub.shape[0], ip(dyn), ip(par), dp(diag), dp(lo), dp(up), ip(edges),
                                     ip(lvl), ip(jumps), dp(arrs[8]), dp(g), dp(camask), dp(clamp), on, off, int(scv),
                                     dp(V), dp(S))
            if n == 0:
                p = o.params
                self.const = PortConst(
                    o.E_Na, o.E_K, float(p["IL"].get("E", -70.0)), float(p["Na_Ba2002"].get("V_sh", -50.0)),
                    o._phi("Na_Ba2002"), float(p["KDR_Ba2002"].get("V_sh", -50.0)), o._phi("KDR_Ba2002"),
                    float(p["CaL_IS2008"].get("V_sh", 0.0)), o._phi("CaL_IS2008", "p"), o._phi("CaL_IS2008", "q"),
                    10.0 / (2.0 * ch.FARADAY * float(o.ca["d"])), 1.0 / float(o.ca["tau"]), float(o.ca["C_rest"]),
                    float(o.ca.get("Co", 2.0)), 1e3 * ch.GAS_CONSTANT * float(o.ca["temp"]) / (2.0 * ch.FARADAY))
        self.n = len(models)
        self.n_cv = sum(c.C for c in self.cells)

    def run(self, n_steps, step0=0, sequential=True, threads=0):
        L = _lib()
        tr = np.zeros((n_steps, self.n))
        rc = L.port_run(self.n, self.cells, ctypes.byref(self.const), self.dt, step0, n_steps,
                        tr.ctypes.data_as(D), int(sequential), int(threads))
        if rc != 0:
            raise MemoryError("port_run workspace allocation failed")
        return tr

    @staticmethod
    def max_threads():
        return _lib().port_max_threads()
