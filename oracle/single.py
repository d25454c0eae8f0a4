"""Oracle restatement (FP64) of a SingleCompartment population integrated with
``ind_exp_euler``.  Test infrastructure only.

``SingleCompartment.update`` (``braincell/_single_compartment/base.py:252-278``)
calls ``ind_exp_euler_step`` (``braincell/quad/_exp_euler.py:217-320, 330-478``):
for every DiffEqState y_k (V, every gate, Ci) it evaluates the derivative f_k
and its self-Jacobian lam_k = d f_k / d y_k with *all* states at their old
values (Jacobi), then sets y_k <- y_k + dt exprel(dt lam_k) f_k.

- dV/dt = (I_ext + sum_ch current(V)) / C  (``base.py:212-233``); I_ext is a
  density (nA -> uA/cm^2 through ``area = 2 pi r L``, ``base.py:150-152``).
  lam_V = -sum(g_eff)/C (no channel current depends on V through anything but
  the driving force).
- gates: HH template (``channel/_base.py:575-588``), lam = -phi (a + b) or
  -phi / tau; AHP's alpha reads Ci (``potassium_calcium.py:179-185``).
- Ci (CalciumDetailed, ``ion/calcium.py:499-503``) uses the ion's own current
  evaluated at the old state (no step-start cache on this path,
  ``ion/_base.py:490-506``): f = max(I_Ca / (2 F d), 0) + (C_rest - Ci)/tau,
  lam = -1/tau + [drive > 0] d(drive)/dCi with dE_Ca/dCi = -RT/(zF Ci).

The thalamocortical template is the reference README's HTC example
(``README.md:52-71``).

``implicit_step(dt, theta)`` restates the linearly implicit solvers on the
same 11-state vector y = (V, nine gates, Ci):

- ``backward_euler`` (``quad/_backward_euler.py:31-64``): J = df/dy at y_n
  (``jacrev_last_dim``, exact derivatives), (I - dt J) dy = dt f(y_n);
- ``implicit_euler`` (``quad/_implicit.py:43-120``, order 2): the Newton
  iterate from y_n of y - y_n - dt/2 (f(y_n) + f(y)) = 0, i.e.
  (I - dt/2 J) dy = dt f(y_n).  Further iterates happen in the reference
  only while the population-wide residual before an update is below its
  tol = 1e-5 and are not restated.

J is taken by complex-step differentiation of ``derivatives`` (exact to
rounding), and each neuron's 11 x 11 system is solved densely.
"""

from __future__ import annotations

import numpy as np

from . import channels as ch

NAMES = ("Na_Ba2002", "K_Leak", "KDR_Ba2002", "CaL_IS2008", "CaT_HM1992", "AHP_De1994", "HCN_HM1992", "IL")


class HTCPopulationOracle:
    """FP64 Jacobi exp-Euler over N thalamocortical point neurons.

    ``g``: dict name -> [N] densities in S/cm^2; ``I_dc_nA``: [N];
    ``params``: per-channel non-density parameters; ``ions``: E_Na, E_K and
    the Ca pool parameters.
    """

    def __init__(self, g, I_dc_nA, params, ions, length_um=10.0, radius_um=5.0, C=1.0):
        self.g = {k: np.asarray(v, dtype=np.float64) * 1e3 for k, v in g.items()}  # mS/cm^2
        self.N = len(next(iter(self.g.values())))
        area_cm2 = 2.0 * np.pi * radius_um * length_um * 1e-8
        self.I_ext = np.asarray(I_dc_nA, dtype=np.float64) * 1e-3 / area_cm2  # uA/cm^2
        self.p = params
        self.E_Na = float(ions["E_Na"])
        self.E_K = float(ions["E_K"])
        ca = ions["ca"]
        self.ca = ca
        self.C = float(C)
        na, kd = params["Na_Ba2002"], params["KDR_Ba2002"]
        self.na_vsh = na.get("V_sh", -50.0)
        self.na_phi = ch.q10_phi(na.get("q10", 3.0), na.get("temp", ch.c2k(36)), na.get("temp_ref", ch.c2k(36)))
        self.kd_vsh = kd.get("V_sh", -50.0)
        self.kd_phi = ch.q10_phi(kd.get("q10", 3.0), kd.get("temp", ch.c2k(36)), kd.get("temp_ref", ch.c2k(36)))
        cl, ct = params["CaL_IS2008"], params["CaT_HM1992"]
        self.cl_vsh = cl.get("V_sh", 0.0)
        self.cl_phi = (ch.q10_phi(cl.get("q10_p", 3.55), cl.get("temp", ch.c2k(36)), cl.get("temp_ref_p", ch.c2k(24))),
                       ch.q10_phi(cl.get("q10_q", 3.0), cl.get("temp", ch.c2k(36)), cl.get("temp_ref_q", ch.c2k(24))))
        self.ct_vsh = ct.get("V_sh", -3.0)
        self.ct_phi = (ch.q10_phi(ct.get("q10_p", 3.55), ct.get("temp", ch.c2k(36)), ct.get("temp_ref_p", ch.c2k(24))),
                       ch.q10_phi(ct.get("q10_q", 3.0), ct.get("temp", ch.c2k(36)), ct.get("temp_ref_q", ch.c2k(24))))
        ah = params["AHP_De1994"]
        self.ahp = (ah.get("alpha", 48.0), ah.get("beta", 0.09), ah.get("n", 2.0), ah.get("phi", 1.0))
        hc = params["HCN_HM1992"]
        self.hcn_E = hc.get("E", 43.0)
        self.hcn_phi = ch.q10_phi(hc.get("q10", 1.0), hc.get("temp", ch.c2k(36)), hc.get("temp_ref", ch.c2k(36)))
        self.il_E = params["IL"].get("E", -70.0)

    def reset_state(self, V0):
        """V at V0; gates at steady state; Ci at its initializer."""
        V = np.asarray(V0, dtype=np.float64).copy()
        self.V = V
        ap, bp, aq, bq = ch.na_ba2002(V, self.na_vsh)
        a, b = ch.kdr_ba2002(V, self.kd_vsh)
        cpi, _, cqi, _ = ch.cal_is2008(V, self.cl_vsh)
        tpi, _, tqi, _ = ch.cat_hm1992(V, self.ct_vsh)
        self.Ci = np.full(self.N, float(self.ca["Ci0"]))
        aa, ab = ch.ahp_de1994(self.Ci, *self.ahp[:3])
        hpi, _ = ch.hcn_hm1992(V)
        self.s = {"na_p": ap / (ap + bp), "na_q": aq / (aq + bq), "kd_p": a / (a + b), "cl_p": cpi, "cl_q": cqi,
                  "ct_p": tpi, "ct_q": tqi, "ahp_p": aa / (aa + ab), "hcn_p": hpi}

    def _E_ca(self, Ci):
        return ch.nernst(Ci, float(self.ca.get("Co", 2.0)), float(self.ca.get("temp", ch.c2k(36))), 2.0)

    def conductances(self, s):
        g = self.g
        return {
            "Na_Ba2002": g["Na_Ba2002"] * s["na_p"] ** 3 * s["na_q"], "K_Leak": g["K_Leak"],
            "KDR_Ba2002": g["KDR_Ba2002"] * s["kd_p"] ** 4, "CaL_IS2008": g["CaL_IS2008"] * s["cl_p"] ** 2 * s["cl_q"],
            "CaT_HM1992": g["CaT_HM1992"] * s["ct_p"] ** 2 * s["ct_q"], "AHP_De1994": g["AHP_De1994"] * s["ahp_p"] ** 2,
            "HCN_HM1992": g["HCN_HM1992"] * s["hcn_p"], "IL": g["IL"]}

    def derivatives(self, V, s, Ci):
        """Returns dict of (f, lam) per state, all at the given (old) values."""
        ge = self.conductances(s)
        Eca = self._E_ca(Ci)
        E = {"Na_Ba2002": self.E_Na, "K_Leak": self.E_K, "KDR_Ba2002": self.E_K, "CaL_IS2008": Eca,
             "CaT_HM1992": Eca, "AHP_De1994": self.E_K, "HCN_HM1992": self.hcn_E, "IL": self.il_E}
        I = self.I_ext + sum(ge[n] * (E[n] - V) for n in NAMES)
        G = sum(ge[n] for n in NAMES)
        out = {"V": (I / self.C, -G / self.C)}
        ap, bp, aq, bq = ch.na_ba2002(V, self.na_vsh)
        out["na_p"] = self._ab(s["na_p"], ap, bp, self.na_phi)
        out["na_q"] = self._ab(s["na_q"], aq, bq, self.na_phi)
        a, b = ch.kdr_ba2002(V, self.kd_vsh)
        out["kd_p"] = self._ab(s["kd_p"], a, b, self.kd_phi)
        pi, pt, qi, qt = ch.cal_is2008(V, self.cl_vsh)
        out["cl_p"] = self._it(s["cl_p"], pi, pt, self.cl_phi[0])
        out["cl_q"] = self._it(s["cl_q"], qi, qt, self.cl_phi[1])
        pi, pt, qi, qt = ch.cat_hm1992(V, self.ct_vsh)
        out["ct_p"] = self._it(s["ct_p"], pi, pt, self.ct_phi[0])
        out["ct_q"] = self._it(s["ct_q"], qi, qt, self.ct_phi[1])
        aa, ab = ch.ahp_de1994(Ci, *self.ahp[:3])
        out["ahp_p"] = self._ab(s["ahp_p"], aa, ab, self.ahp[3])
        pi, pt = ch.hcn_hm1992(V)
        out["hcn_p"] = self._it(s["hcn_p"], pi, pt, self.hcn_phi)
        gca = ge["CaL_IS2008"] + ge["CaT_HM1992"]
        I_ca = gca * (Eca - V)
        dfac = 10.0 / (2.0 * ch.FARADAY * float(self.ca["d"]))
        raw = I_ca * dfac
        drive = np.maximum(raw, 0.0)
        tau = float(self.ca["tau"])
        f = drive + (float(self.ca["C_rest"]) - Ci) / tau
        rtzf = ch.GAS_CONSTANT * float(self.ca.get("temp", ch.c2k(36))) / (2.0 * ch.FARADAY) * 1e3
        dE = -rtzf / Ci
        lam = -1.0 / tau + np.where(raw > 0.0, dfac * gca * dE, 0.0)
        out["Ci"] = (f, lam)
        return out

    @staticmethod
    def _ab(x, a, b, phi):
        return phi * (a * (1.0 - x) - b * x), -phi * (a + b)

    @staticmethod
    def _it(x, inf, tau, phi):
        return phi * (inf - x) / tau, -phi / tau

    GATES = ("na_p", "na_q", "kd_p", "cl_p", "cl_q", "ct_p", "ct_q", "ahp_p", "hcn_p")

    def state_vector(self):
        """[11, N]: V, the nine gates in kernel slot order, Ci."""
        return np.stack([self.V, *(self.s[k] for k in self.GATES), self.Ci])

    def rhs(self, y):
        """f(y) for y = [11, N] (real or complex)."""
        d = self.derivatives(y[0], dict(zip(self.GATES, y[1:10])), y[10])
        return np.stack([d["V"][0] * np.ones_like(y[0]), *(d[k][0] for k in self.GATES), d["Ci"][0]])

    def jacobian(self, y, h=1e-30):
        """J[i, k, n] = d f_i / d y_k of neuron n, by complex steps of size h."""
        J = np.empty((11, 11, y.shape[1]))
        for k in range(11):
            yc = y.astype(complex)
            yc[k] += 1j * h
            J[:, k] = self.rhs(yc).imag / h
        return J

    def implicit_step(self, dt, theta):
        """theta = 1: backward_euler; theta = 1/2: implicit_euler (order 2)."""
        y = self.state_vector()
        f = self.rhs(y)
        J = self.jacobian(y)
        lhs = np.eye(11)[None] - theta * dt * np.moveaxis(J, 2, 0)
        dy = np.linalg.solve(lhs, (dt * f).T[..., None])[..., 0].T
        y1 = y + dy
        self.V = y1[0]
        self.s = {k: y1[1 + j] for j, k in enumerate(self.GATES)}
        self.Ci = y1[10]

    def step(self, dt):
        d = self.derivatives(self.V, self.s, self.Ci)
        Vn = ch.exp_euler(self.V, *d["V"], dt)
        sn = {k: ch.exp_euler(v, *d[k], dt) for k, v in self.s.items()}
        Cin = ch.exp_euler(self.Ci, *d["Ci"], dt)
        self.V, self.s, self.Ci = Vn, sn, Cin

    def run(self, dt, n_steps, probes=None, theta=None):
        """theta None: ind_exp_euler; 1 or 1/2: ``implicit_step``."""
        probes = np.arange(self.N) if probes is None else np.asarray(probes)
        out = np.empty((n_steps, len(probes)))
        for n in range(n_steps):
            if theta is None:
                self.step(dt)
            else:
                self.implicit_step(dt, theta)
            out[n] = self.V[probes]
        return out
