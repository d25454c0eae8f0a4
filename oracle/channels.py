"""Oracle restatement (FP64) of the channel and ion mechanism math on the
staggered path.  Test infrastructure only.

Units: V in mV, rates in 1/ms, tau in ms, g in mS/cm^2 (currents in uA/cm^2),
Ci/Co in mM, temperatures in K.

Reference sources, per function below:

- HH template: ``braincell/channel/_base.py:517-532`` (gate_phi),
  ``:550-558`` (conductance_factor), ``:560-573`` (reset_state),
  ``:575-588`` (compute_derivative), ``:658-674`` (OhmicHH.current)
- ``Na_Ba2002`` ``channel/sodium.py:134-170``
- ``Na_HH1952`` ``channel/sodium.py:293-398``
- ``K_HH1952`` ``channel/potassium.py:341-456``
- ``KDR_Ba2002`` ``channel/potassium.py:189-214``
- ``K_Leak`` ``channel/potassium.py:1216-1235``
- ``IL`` ``channel/leaky.py:184-200``
- ``CaT_HM1992`` ``channel/calcium.py:329-373``
- ``CaL_IS2008`` ``channel/calcium.py:880-920``
- ``AHP_De1994`` ``channel/potassium_calcium.py:160-185``
- ``HCN_HM1992`` ``channel/hyperpolarization_activated.py:103-132``
- ``CalciumDetailed.derivative`` ``ion/calcium.py:499-503`` and the Nernst
  potential ``ion/_base.py:462-469``
- ind_exp_euler update ``quad/_exp_euler.py:437-475``

``exprel`` is ``scipy.special.exprel`` (brainunit's ``u.math.exprel`` is the
same function, (e^x - 1)/x with value 1 at 0; its float behaviour near 0 is
parity unpinned).  F and R are CODATA 2018.
"""

from __future__ import annotations

import numpy as np
from scipy.special import exprel

FARADAY = 96485.33212  # C/mol (CODATA 2018)
GAS_CONSTANT = 8.314462618  # J/(mol K) (CODATA 2018)


def _exprel(x):
    """``exprel`` for real input; complex input (complex-step derivatives of
    the rates) takes expm1(x)/x, with the Taylor series where |x| < 1e-6."""
    if not np.iscomplexobj(x):
        return exprel(x)
    x = np.asarray(x)
    small = np.abs(x) < 1e-6
    xs = np.where(small, 1.0, x)
    return np.where(small, 1.0 + x * (0.5 + x / 6.0), np.expm1(xs) / xs)


def c2k(celsius: float) -> float:
    return float(celsius) + 273.15


def q10_phi(q10, temp_k, temp_ref_k):
    """_base.py:531: q10 ** ((T - T_ref) / 10 K)."""
    return q10 ** ((temp_k - temp_ref_k) / 10.0)


# ---------------------------------------------------------------- rate tables
def na_ba2002(V, V_sh=-50.0):
    """sodium.py:158-170 -> (alpha_p, beta_p, alpha_q, beta_q)."""
    Vp = V - V_sh
    t = Vp - 13.0
    ap = 0.32 * 4.0 / _exprel(-t / 4.0)
    t = Vp - 40.0
    bp = 0.28 * 5.0 / _exprel(t / 5.0)
    aq = 0.128 * np.exp(-(Vp - 17.0) / 18.0)
    bq = 4.0 / (1.0 + np.exp(-(Vp - 40.0) / 5.0))
    return ap, bp, aq, bq


def kdr_ba2002(V, V_sh=-50.0):
    """potassium.py:206-214 -> (alpha_p, beta_p)."""
    Vp = V - V_sh
    t = Vp - 15.0
    ap = 0.032 * 5.0 / _exprel(-t / 5.0)
    bp = 0.5 * np.exp(-(Vp - 10.0) / 40.0)
    return ap, bp


def na_hh1952(V, V_sh=-45.0):
    """sodium.py:384-398 -> (alpha_p, beta_p, alpha_q, beta_q)."""
    Vp = V - V_sh
    ap = 1.0 / _exprel(-(Vp - 5.0) / 10.0)
    bp = 4.0 * np.exp(-(Vp + 20.0) / 18.0)
    aq = 0.07 * np.exp(-(Vp + 20.0) / 20.0)
    bq = 1.0 / (1.0 + np.exp(-(Vp - 10.0) / 10.0))
    return ap, bp, aq, bq


def k_hh1952(V, V_sh=-45.0):
    """potassium.py:450-456 -> (alpha_p, beta_p)."""
    Vp = V - V_sh
    ap = 0.1 / _exprel(-(Vp + 10.0) / 10.0)
    bp = 0.125 * np.exp(-(Vp + 20.0) / 80.0)
    return ap, bp


def cat_hm1992(V, V_sh=-3.0):
    """calcium.py:345-373 -> (p_inf, p_tau, q_inf, q_tau)."""
    Vp = V - V_sh
    p_inf = 1.0 / (1.0 + np.exp(-(Vp + 59.0) / 6.2))
    p_tau = 1.0 / (np.exp(-(Vp + 132.0) / 16.7) + np.exp((Vp + 16.8) / 18.2)) + 0.612
    q_inf = 1.0 / (1.0 + np.exp((Vp + 83.0) / 4.0))
    q_tau = np.where(Vp >= -80.0, np.exp(-(Vp + 22.0) / 10.5) + 28.0, np.exp((Vp + 467.0) / 66.6))
    return p_inf, p_tau, q_inf, q_tau


def cal_is2008(V, V_sh=0.0):
    """calcium.py:902-920 -> (p_inf, p_tau, q_inf, q_tau)."""
    Vp = V - V_sh
    p_inf = 1.0 / (1.0 + np.exp(-(Vp + 10.0) / 4.0))
    p_tau = 0.4 + 0.7 / (np.exp(-(Vp + 5.0) / 15.0) + np.exp((Vp + 5.0) / 15.0))
    q_inf = 1.0 / (1.0 + np.exp((Vp + 25.0) / 2.0))
    q_tau = 300.0 + 100.0 / (np.exp((Vp + 40.0) / 9.5) + np.exp(-(Vp + 40.0) / 9.5))
    return p_inf, p_tau, q_inf, q_tau


def ahp_de1994(Ci, alpha=48.0, beta=0.09, n=2.0):
    """potassium_calcium.py:179-185 -> (alpha_p, beta_p)."""
    return alpha * np.power(Ci, n), beta * np.ones(np.shape(Ci))


def hcn_hm1992(V):
    """hyperpolarization_activated.py:124-132 -> (p_inf, p_tau)."""
    p_inf = 1.0 / (1.0 + np.exp((V + 75.0) / 5.5))
    p_tau = 1.0 / (np.exp(-0.086 * V - 14.59) + np.exp(0.0701 * V - 1.87))
    return p_inf, p_tau


def nernst(Ci, Co=2.0, temp_k=c2k(36.0), valence=2.0):
    """ion/_base.py:462-469: RT/(zF) ln(Co/Ci) in mV."""
    return GAS_CONSTANT * temp_k / (valence * FARADAY) * np.log(Co / Ci) * 1e3


def ca_detailed_drive(I_ca, d_um):
    """ion/calcium.py:499-503 drive term in mM/ms: max(I / (2 F d), 0).

    I in uA/cm^2 (1e-2 A/m^2), d in um: I[A/m^2] / (2 F d[m]) is mol/(m^3 s)
    = mM/s, and /1e3 gives mM/ms, i.e. a factor 10 overall.
    """
    drive = I_ca * 10.0 / (2.0 * FARADAY * d_um)
    return np.maximum(drive, 0.0)


# ------------------------------------------------------------- exp-Euler core
def exp_euler(y, f, lam, dt):
    """_exp_euler.py:462-465: y + dt * exprel(dt * lam) * f."""
    return y + dt * exprel(dt * lam) * f


def ab_update(x, a, b, phi, dt):
    """Gate in alpha/beta form: f = phi (a (1-x) - b x), lam = -phi (a+b)."""
    f = phi * (a * (1.0 - x) - b * x)
    lam = -phi * (a + b)
    return exp_euler(x, f, lam, dt)


def it_update(x, inf, tau, phi, dt):
    """Gate in inf/tau form: f = phi (inf - x)/tau, lam = -phi / tau."""
    f = phi * (inf - x) / tau
    lam = -phi / tau
    return exp_euler(x, f, lam, dt)
