"""FP64 reverse pass of the staggered step on the oracle (test infrastructure).

The reference obtains gradients by JAX autodiff through its scan; that path
cannot run here, so the check of the hand-written GPU adjoint is (a) central
finite differences of the oracle's FP64 loss and (b) this independent FP64
restatement of the reverse equations over the oracle's own DHS rows.  Model
channel set: Na_Ba2002, KDR_Ba2002, IL, CaL_IS2008 with a CalciumDetailed
pool (the differentiation workload).
"""

from __future__ import annotations

import numpy as np

from . import channels as ch
from .dhs import comp_backsub_raw, comp_triang_raw, numeric_state
from .staggered import StaggeredOracle

NAMES = ("Na_Ba2002", "KDR_Ba2002", "IL", "CaL_IS2008")


def _rate_d(fun, V, *args, h=1e-6):
    """Central difference of a rate function (used only for d/dV of rates)."""
    return [(p - m) / (2 * h) for p, m in zip(fun(V + h, *args), fun(V - h, *args))]


def forward_states(o: StaggeredOracle, dt, n_steps):
    """Run the oracle storing (V, gates, Ci) at every step start."""
    states = []
    for n in range(n_steps):
        states.append((o.V.copy(), {k: [g.copy() for g in v] for k, v in o.gates.items()}, o.Ci.copy()))
        o.step(n * dt, dt)
    states.append((o.V.copy(), {k: [g.copy() for g in v] for k, v in o.gates.items()}, o.Ci.copy()))
    return states


def gate_grads(x, a, b, da, db, phi, dt):
    u = -dt * phi * (a + b)
    er = ch.exprel(u)
    h = 1e-6
    erd = (ch.exprel(u + h) - ch.exprel(u - h)) / (2 * h)
    f = phi * (a * (1 - x) - b * x)
    common = dt * erd * (-dt * phi) * f
    return np.exp(u), (common + dt * er * phi * (1 - x)) * da + (common - dt * er * phi * x) * db


def gate_grads_it(x, inf, tau, dinf, dtau, phi, dt):
    w = 1.0 / tau
    u = -dt * phi * w
    er = ch.exprel(u)
    h = 1e-6
    erd = (ch.exprel(u + h) - ch.exprel(u - h)) / (2 * h)
    f = phi * (inf - x) * w
    dI = dt * er * phi * w
    dW = dt * erd * (-dt * phi) * f + dt * er * phi * (inf - x)
    return np.exp(u), dI * dinf + dW * (-dtau * w * w)


def adjoint(o: StaggeredOracle, states, dt, trace_bar, probe_cv):
    """Gradient of sum_n trace_bar[n] * V_{n+1}[probe_cv] w.r.t. every
    channel's per-CV density (mS/cm^2).  Returns dict name -> [C]."""
    C = o.C
    st = o.st
    ca = o.ca
    cm = o.cm
    n_steps = len(states) - 1
    Vb = np.zeros(C)
    gb = {"na_p": np.zeros(C), "na_q": np.zeros(C), "kd_p": np.zeros(C), "cl_p": np.zeros(C), "cl_q": np.zeros(C)}
    cib = np.zeros(C)
    grad = {n: np.zeros(C) for n in NAMES}
    g = o.g
    tau = float(ca["tau"])
    kdrive = 10.0 / (2.0 * ch.FARADAY * float(ca["d"]))
    rtzf = ch.GAS_CONSTANT * float(ca["temp"]) / (2.0 * ch.FARADAY) * 1e3
    na_vsh = float(o.params["Na_Ba2002"].get("V_sh", -50.0))
    kd_vsh = float(o.params["KDR_Ba2002"].get("V_sh", -50.0))
    cl_vsh = float(o.params["CaL_IS2008"].get("V_sh", 0.0))
    E_IL = float(o.params["IL"].get("E", -70.0))
    for n in range(n_steps - 1, -1, -1):
        V0, G0, Ci0 = states[n]
        X = states[n + 1][0]
        na_p, na_q = G0["Na_Ba2002"]
        kd_p = G0["KDR_Ba2002"][0]
        cl_p, cl_q = G0["CaL_IS2008"]
        Vb[probe_cv] += trace_bar[n]
        Xb = Vb.copy()
        # gates at X
        ap, bp, aq, bq = ch.na_ba2002(X, na_vsh)
        dap, dbp, daq, dbq = _rate_d(ch.na_ba2002, X, na_vsh)
        e, dv = gate_grads(na_p, ap, bp, dap, dbp, o._phi("Na_Ba2002"), dt)
        Xb += gb["na_p"] * dv
        nb_p = gb["na_p"] * e
        e, dv = gate_grads(na_q, aq, bq, daq, dbq, o._phi("Na_Ba2002"), dt)
        Xb += gb["na_q"] * dv
        nb_q = gb["na_q"] * e
        a, b = ch.kdr_ba2002(X, kd_vsh)
        da, db = _rate_d(ch.kdr_ba2002, X, kd_vsh)
        e, dv = gate_grads(kd_p, a, b, da, db, o._phi("KDR_Ba2002"), dt)
        Xb += gb["kd_p"] * dv
        nk_p = gb["kd_p"] * e
        pi, pt, qi, qt = ch.cal_is2008(X, cl_vsh)
        dpi, dpt, dqi, dqt = _rate_d(ch.cal_is2008, X, cl_vsh)
        e, dv = gate_grads_it(cl_p, pi, pt, dpi, dpt, o._phi("CaL_IS2008", "p"), dt)
        Xb += gb["cl_p"] * dv
        nc_p = gb["cl_p"] * e
        e, dv = gate_grads_it(cl_q, qi, qt, dqi, dqt, o._phi("CaL_IS2008", "q"), dt)
        Xb += gb["cl_q"] * dv
        nc_q = gb["cl_q"] * e
        # Ca pool
        u = -dt / tau
        ec = np.exp(u)
        bcf = dt * ch.exprel(u)
        eca = ch.nernst(Ci0, float(ca["Co"]), float(ca["temp"]), 2.0)
        gcal = g["CaL_IS2008"] * cl_p ** 2 * cl_q
        mask = o.ca_mask.astype(float)
        raw = mask * gcal * (eca - V0) * kdrive
        Ib = np.where(raw > 0, cib * bcf * kdrive * mask, 0.0)
        Vb_new = -Ib * gcal
        cin = cib * ec + Ib * gcal * (-rtzf / Ci0)
        drv = Ib * (eca - V0)
        nc_p = nc_p + drv * g["CaL_IS2008"] * 2 * cl_p * cl_q
        nc_q = nc_q + drv * g["CaL_IS2008"] * cl_p ** 2
        grad["CaL_IS2008"] += drv * cl_p ** 2 * cl_q
        # transposed solve with the step's matrix
        G = g["Na_Ba2002"] * na_p ** 3 * na_q + g["KDR_Ba2002"] * kd_p ** 4 + g["IL"] + gcal
        lin = -G / cm
        d, s, lo, up = numeric_state(st, V0[None], lin[None], np.zeros((1, C)), dt)
        s[:] = 0.0
        s[0, st.dynamic_rows] = Xb
        dT, sT = comp_triang_raw(d, s, up, lo, st.edges, st.level_offsets)
        y = comp_backsub_raw(dT, sT, up, st.backsub)[0]
        Rb = y[st.dynamic_rows]
        Vb_new = Vb_new + Rb
        w = dt / cm * Rb
        gbar = w * (o.E_Na - X)
        nb_p = nb_p + gbar * g["Na_Ba2002"] * 3 * na_p ** 2 * na_q
        nb_q = nb_q + gbar * g["Na_Ba2002"] * na_p ** 3
        grad["Na_Ba2002"] += gbar * na_p ** 3 * na_q
        gbar = w * (o.E_K - X)
        nk_p = nk_p + gbar * g["KDR_Ba2002"] * 4 * kd_p ** 3
        grad["KDR_Ba2002"] += gbar * kd_p ** 4
        grad["IL"] += w * (E_IL - X)
        cin = cin + w * gcal * (-rtzf / Ci0)
        gbar = w * (eca - X)
        nc_p = nc_p + gbar * g["CaL_IS2008"] * 2 * cl_p * cl_q
        nc_q = nc_q + gbar * g["CaL_IS2008"] * cl_p ** 2
        grad["CaL_IS2008"] += gbar * cl_p ** 2 * cl_q
        Vb = Vb_new
        gb = {"na_p": nb_p, "na_q": nb_q, "kd_p": nk_p, "cl_p": nc_p, "cl_q": nc_q}
        cib = cin
    return grad
