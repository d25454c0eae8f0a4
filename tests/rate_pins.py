"""Characteristic pins of the oracle's rate functions and mechanism math.

Each pin takes a namespace holding the oracle's channel functions
(``oracle.channels`` itself, or a copy of it with one literal changed) and
raises ``AssertionError`` when a characteristic of the published rate law
does not hold.  The characteristics are properties that a rate law's
constants fix, stated in a form other than the oracle's own expression:

* a logistic 1/(1 + e^{-(V - c)/k}) has the value 1/2 at c, slope 1/(4k) there
  and the asymptotes 0 and 1;
* a Traub-Miles / HH linear-exponential rate a u / (1 - e^{-u/k}) with
  u = V - c has the removable value a k at u = 0 and the odd-part identity
  f(c + u) - f(c - u) = a u for every u;
* an exponential rate a e^{-(V - c)/k} equals a at c and falls by e over k;
* a time constant b + A / (e^{-(V - c)/k} + e^{(V - c)/k}) peaks at c with
  b + A/2, is symmetric about c and tends to b; a sum of two exponentials in
  its reciprocal is fixed by its log-slopes far from the crossover.

The constants are those of the reference's channel classes
(``braincell/channel/sodium.py:134-170, 384-398``,
``potassium.py:189-214, 440-456``, ``calcium.py:329-373, 880-920``,
``potassium_calcium.py:160-185``, ``hyperpolarization_activated.py:103-132``)
and their documented defaults (V_sh, q10, reference temperatures).
"""

from __future__ import annotations

import math

import numpy as np

E = math.e


def _v(x):
    return np.atleast_1d(np.asarray(x, dtype=float))


def _eq(got, want, rel=1e-9, abs_=0.0, what=""):
    got = np.asarray(got, dtype=float)
    assert np.allclose(got, want, rtol=rel, atol=abs_), f"{what}: {got} != {want}"


def _slope(f, x, h=1e-4):
    return (f(x + h) - f(x - h)) / (2 * h)


# ------------------------------------------------------------------ shapes
def logistic_pin(f, c, k, what, top=1.0):
    """f = top / (1 + e^{-(V - c)/k}) (k < 0 for a falling logistic)."""
    _eq(f(_v(c)), top / 2, what=f"{what} midpoint")
    _eq(_slope(f, _v(c)), top / (4 * k), rel=1e-7, what=f"{what} midpoint slope")
    hi = c + 60.0 * abs(k) * np.sign(k)
    lo = c - 60.0 * abs(k) * np.sign(k)
    _eq(f(_v(hi)), top, rel=1e-12, what=f"{what} upper asymptote")
    _eq(f(_v(lo)), 0.0, abs_=1e-20, what=f"{what} lower asymptote")


def linexp_pin(f, a, c, k, what):
    """f = a u / (1 - e^{-u/k}), u = s (V - c) with s = sign(k) folded into k:
    the value a |k| at V = c and the odd-part identity."""
    _eq(f(_v(c)), a * abs(k), rel=1e-12, what=f"{what} removable value")
    for u in (0.7, 3.0, 11.0, 37.0):
        s = 1.0 if k > 0 else -1.0
        _eq(f(_v(c + s * u)) - f(_v(c - s * u)), a * u, rel=1e-10, what=f"{what} odd part at u={u}")


def expo_pin(f, a, c, k, what):
    """f = a e^{-(V - c)/k}."""
    _eq(f(_v(c)), a, rel=1e-12, what=f"{what} value at {c}")
    _eq(f(_v(c - k)) / f(_v(c)), E, rel=1e-12, what=f"{what} e-fold over {k} mV")


def coshtau_pin(f, base, amp, c, k, what):
    """f = base + amp / (e^{-(V - c)/k} + e^{(V - c)/k})."""
    _eq(f(_v(c)), base + amp / 2, rel=1e-12, what=f"{what} peak")
    for u in (1.0, 7.0, 23.0):
        _eq(f(_v(c + u)), f(_v(c - u)), rel=1e-12, what=f"{what} symmetry at {u}")
    _eq(f(_v(c + k)), base + amp / (2 * math.cosh(1.0)), rel=1e-12, what=f"{what} one slope from the peak")
    _eq(f(_v(c + 60 * k)), base, rel=1e-12, what=f"{what} floor")


def two_exp_rate_pin(g, lo_pts, hi_pts, lo_law, hi_law, what):
    """g = 1 / (tau - floor) is a sum of two exponentials; far below the
    crossover log g follows lo_law (slope, intercept), far above hi_law."""
    for pts, (sl, ic), side in ((lo_pts, lo_law, "low"), (hi_pts, hi_law, "high")):
        for v in pts:
            _eq(np.log(g(_v(v))), sl * v + ic, rel=1e-9, abs_=1e-9, what=f"{what} {side}-side log law at {v}")


# ---------------------------------------------------------------- channels
def pin_na_ba2002(m):
    vsh = -50.0  # sodium.py:144 default

    def r(i):
        return lambda V: m.na_ba2002(V)[i]
    # alpha_p = 0.32 (13 - V') / (e^{(13 - V')/4} - 1), rising in V
    linexp_pin(r(0), 0.32, vsh + 13.0, 4.0, "Na_Ba2002 alpha_p")
    # beta_p = 0.28 (V' - 40) / (e^{(V' - 40)/5} - 1): falls with V, so the
    # odd-part identity is f(c - u) - f(c + u) = 0.28 u
    f = r(1)
    c = vsh + 40.0
    _eq(f(_v(c)), 0.28 * 5.0, rel=1e-12, what="Na_Ba2002 beta_p removable value")
    for u in (0.7, 3.0, 11.0, 37.0):
        _eq(f(_v(c - u)) - f(_v(c + u)), 0.28 * u, rel=1e-10, what=f"Na_Ba2002 beta_p odd part at {u}")
    expo_pin(r(2), 0.128, vsh + 17.0, 18.0, "Na_Ba2002 alpha_q")
    logistic_pin(r(3), vsh + 40.0, 5.0, "Na_Ba2002 beta_q", top=4.0)


def pin_kdr_ba2002(m):
    vsh = -50.0  # potassium.py:199 default
    f = lambda V: m.kdr_ba2002(V)[0]
    # alpha_p = 0.032 (15 - V') / (e^{(15 - V')/5} - 1)
    linexp_pin(f, 0.032, vsh + 15.0, 5.0, "KDR_Ba2002 alpha_p")
    expo_pin(lambda V: m.kdr_ba2002(V)[1], 0.5, vsh + 10.0, 40.0, "KDR_Ba2002 beta_p")


def pin_na_hh1952(m):
    vsh = -45.0  # sodium.py:374 default
    r = lambda i: (lambda V: m.na_hh1952(V)[i])
    # HH 1952: alpha_m = 0.1 (V + 40) / (1 - e^{-(V + 40)/10}) for V_sh = -45
    linexp_pin(r(0), 0.1, vsh + 5.0, 10.0, "Na_HH1952 alpha_p")
    expo_pin(r(1), 4.0, vsh - 20.0, 18.0, "Na_HH1952 beta_p")
    expo_pin(r(2), 0.07, vsh - 20.0, 20.0, "Na_HH1952 alpha_q")
    logistic_pin(r(3), vsh + 10.0, 10.0, "Na_HH1952 beta_q")


def pin_k_hh1952(m):
    vsh = -45.0  # potassium.py:440 default
    linexp_pin(lambda V: m.k_hh1952(V)[0], 0.01, vsh - 10.0, 10.0, "K_HH1952 alpha_p")
    expo_pin(lambda V: m.k_hh1952(V)[1], 0.125, vsh - 20.0, 80.0, "K_HH1952 beta_p")


def pin_cat_hm1992(m):
    vsh = -3.0  # calcium.py:340 default
    r = lambda i: (lambda V: m.cat_hm1992(V)[i])
    logistic_pin(r(0), vsh - 59.0, 6.2, "CaT_HM1992 p_inf")
    # tau_p = 1 / (e^{-(V' + 132)/16.7} + e^{(V' + 16.8)/18.2}) + 0.612
    _eq(r(1)(_v(vsh + 400.0)), 0.612, rel=1e-9, what="CaT_HM1992 tau_p floor")
    g = lambda V: 1.0 / (r(1)(V) - 0.612)
    two_exp_rate_pin(g, [vsh - 400.0, vsh - 383.3], [vsh + 150.0, vsh + 168.2],
                     (-1 / 16.7, -(132.0 - vsh) / 16.7), (1 / 18.2, (16.8 - vsh) / 18.2), "CaT_HM1992 tau_p")
    logistic_pin(r(2), vsh - 83.0, -4.0, "CaT_HM1992 q_inf")
    q = r(3)
    # V' >= -80: e^{-(V' + 22)/10.5} + 28
    _eq(q(_v(vsh - 22.0)), 29.0, rel=1e-12, what="CaT_HM1992 tau_q upper branch at -22")
    _eq((q(_v(vsh - 32.5)) - 28.0) / (q(_v(vsh - 22.0)) - 28.0), E, rel=1e-12, what="CaT_HM1992 tau_q e-fold 10.5")
    _eq(q(_v(vsh + 300.0)), 28.0, rel=1e-12, what="CaT_HM1992 tau_q floor 28")
    # V' < -80: e^{(V' + 467)/66.6}
    _eq(q(_v(vsh - 467.0)), 1.0, rel=1e-12, what="CaT_HM1992 tau_q lower branch at -467")
    _eq(q(_v(vsh - 400.4)) / q(_v(vsh - 467.0)), E, rel=1e-12, what="CaT_HM1992 tau_q e-fold 66.6")
    # the branch point sits at V' = -80 (upper branch includes it)
    _eq(q(_v(vsh - 80.0)), math.exp(58.0 / 10.5) + 28.0, rel=1e-12, what="CaT_HM1992 tau_q at the branch point")
    _eq(q(_v(vsh - 80.5)), math.exp(386.5 / 66.6), rel=1e-12, what="CaT_HM1992 tau_q just below the branch point")
    _eq(q(_v(vsh - 73.0)), math.exp(51.0 / 10.5) + 28.0, rel=1e-12, what="CaT_HM1992 tau_q above the branch point")
    _eq(q(_v(vsh - 87.0)), math.exp(380.0 / 66.6), rel=1e-12, what="CaT_HM1992 tau_q below the branch point")


def pin_cal_is2008(m):
    vsh = 0.0  # calcium.py:889 default
    r = lambda i: (lambda V: m.cal_is2008(V)[i])
    logistic_pin(r(0), vsh - 10.0, 4.0, "CaL_IS2008 p_inf")
    coshtau_pin(r(1), 0.4, 0.7, vsh - 5.0, 15.0, "CaL_IS2008 tau_p")
    logistic_pin(r(2), vsh - 25.0, -2.0, "CaL_IS2008 q_inf")
    coshtau_pin(r(3), 300.0, 100.0, vsh - 40.0, 9.5, "CaL_IS2008 tau_q")


def pin_ahp_de1994(m):
    a, b = m.ahp_de1994(_v([1.0, 2.0, 0.5]))
    _eq(a, [48.0, 192.0, 12.0], rel=1e-12, what="AHP_De1994 alpha = 48 Ci^2")
    _eq(b, [0.09, 0.09, 0.09], rel=1e-12, what="AHP_De1994 beta = 0.09")


def pin_hcn_hm1992(m):
    logistic_pin(lambda V: m.hcn_hm1992(V)[0], -75.0, -5.5, "HCN_HM1992 p_inf")
    g = lambda V: 1.0 / m.hcn_hm1992(V)[1]
    # 1/tau = e^{-0.086 V - 14.59} + e^{0.0701 V - 1.87}
    two_exp_rate_pin(g, [-400.0, -300.0], [300.0, 400.0], (-0.086, -14.59), (0.0701, -1.87), "HCN_HM1992 tau_p")


# --------------------------------------------------------- ions and updates
def pin_physics(m):
    import scipy.constants as sc

    F = sc.physical_constants["Faraday constant"][0]
    _eq(m.FARADAY, F, rel=1e-10, what="Faraday constant (CODATA 2018)")
    _eq(m.GAS_CONSTANT, sc.R, rel=1e-10, what="gas constant (CODATA 2018)")
    _eq(m.c2k(0.0), sc.zero_Celsius, rel=1e-15, what="0 degC in K")
    # z F E = R T ln(Co / Ci), with scipy's constants
    T = 309.15
    _eq(m.nernst(2.4e-4, 2.0, T, 2.0), 1e3 * sc.R * T / (2 * F) * np.log(2.0 / 2.4e-4), rel=1e-10, what="Nernst")
    _eq(m.nernst(2.0, 2.0, T, 2.0), 0.0, abs_=0.0, what="Nernst at Co = Ci")
    # defaults: Co = 2 mM, T = 36 degC, z = 2 (ion/calcium.py:116-122)
    _eq(m.nernst(2.4e-4), 1e3 * sc.R * (36.0 + sc.zero_Celsius) / (2 * F) * np.log(2.0 / 2.4e-4), rel=1e-10,
        what="Nernst defaults")
    # 1 uA/cm^2 into a 1 um shell: 1e-2 A/m^2 / (2 F 1e-6 m) mol/(m^3 s) = mM/s -> /1e3 per ms
    _eq(m.ca_detailed_drive(_v(1.0), 1.0), 1e-2 / (2 * F * 1e-6) / 1e3, rel=1e-10, what="Ca drive units")
    _eq(m.ca_detailed_drive(_v(-1.0), 1.0), 0.0, abs_=0.0, what="Ca drive rectification")
    # q10 ** ((T - T_ref) / 10 K)
    _eq(m.q10_phi(3.0, 309.15, 299.15), 3.0, rel=1e-14, what="q10 over 10 K")
    _eq(m.q10_phi(3.55, 309.15, 297.15), 3.55 ** 1.2, rel=1e-14, what="q10 over 12 K")


def pin_updates(m):
    # quad/_exp_euler_test.py:113-120: dx/dt = -x, one step of 0.01 gives e^{-0.01}
    _eq(m.exp_euler(1.0, -1.0, -1.0, 0.01), math.exp(-0.01), rel=1e-14, what="exp-Euler linear decay")
    # gate updates equal the exact solution of the linear gate ODE
    x, a, b, phi, dt = 0.3, 0.7, 0.2, 2.0, 0.05
    xi = a / (a + b)
    _eq(m.ab_update(x, a, b, phi, dt), xi + (x - xi) * math.exp(-phi * (a + b) * dt), rel=1e-13,
        what="alpha/beta gate update")
    inf, tau = 0.8, 3.0
    _eq(m.it_update(x, inf, tau, phi, dt), inf + (x - inf) * math.exp(-phi * dt / tau), rel=1e-13,
        what="inf/tau gate update")
    # complex-step exprel near 0: Taylor series 1 + z/2 + z^2/6 + z^3/24
    for z in (1e-7j, 3e-7 + 2e-7j):
        want = 1 + z / 2 + z * z / 6 + z ** 3 / 24
        got = complex(np.asarray(m._exprel(np.asarray([z])))[0])
        assert abs(got - want) < 1e-15, ("complex exprel", z, got, want)
    for z in (1e-3j, 0.4 - 0.3j):
        got = complex(np.asarray(m._exprel(np.asarray([z])))[0])
        want = np.expm1(z) / z
        assert abs(got - want) < 1e-14 * abs(want), ("complex exprel", z, got, want)


PINS = (pin_na_ba2002, pin_kdr_ba2002, pin_na_hh1952, pin_k_hh1952, pin_cat_hm1992, pin_cal_is2008,
        pin_ahp_de1994, pin_hcn_hm1992, pin_physics, pin_updates)
