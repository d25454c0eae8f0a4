"""Host-side (NumPy, FP64) rate functions used only to initialise states.

``reset_state`` in the reference puts every gate at its steady state for the
initial voltage (``braincell/channel/_base.py:560-573``): alpha/(alpha+beta)
for alpha/beta gates, x_inf for inf/tau gates.  The same formulas run on the
GPU inside the kernels (``csrc/bc_channels.h``).
"""

from __future__ import annotations

import numpy as np


def exprel(x):
    x = np.asarray(x, dtype=np.float64)
    small = np.abs(x) < 1e-8
    xs = np.where(small, 1.0, x)
    return np.where(small, 1.0 + x * (0.5 + x / 6.0), np.expm1(xs) / xs)


def na_steady(V, vsh):
    Vp = V - vsh
    ap = 1.28 / exprel(-(Vp - 13.0) / 4.0)
    bp = 1.4 / exprel((Vp - 40.0) / 5.0)
    aq = 0.128 * np.exp(-(Vp - 17.0) / 18.0)
    bq = 4.0 / (1.0 + np.exp(-(Vp - 40.0) / 5.0))
    return ap / (ap + bp), aq / (aq + bq)


def kdr_steady(V, vsh):
    Vp = V - vsh
    ap = 0.16 / exprel(-(Vp - 15.0) / 5.0)
    bp = 0.5 * np.exp(-(Vp - 10.0) / 40.0)
    return ap / (ap + bp)


def cal_steady(V, vsh):
    Vp = V - vsh
    return 1.0 / (1.0 + np.exp(-(Vp + 10.0) / 4.0)), 1.0 / (1.0 + np.exp((Vp + 25.0) / 2.0))


def cat_steady(V, vsh):
    Vp = V - vsh
    return 1.0 / (1.0 + np.exp(-(Vp + 59.0) / 6.2)), 1.0 / (1.0 + np.exp((Vp + 83.0) / 4.0))


def ahp_steady(Ci, alpha, beta, n):
    a = alpha * np.power(Ci, n)
    return a / (a + beta)


def hcn_steady(V):
    return 1.0 / (1.0 + np.exp((V + 75.0) / 5.5))


def nahh_steady(V, vsh):
    """Na_HH1952 (``channel/sodium.py:384-398``)."""
    Vp = V - vsh
    ap = 1.0 / exprel(-(Vp - 5.0) / 10.0)
    bp = 4.0 * np.exp(-(Vp + 20.0) / 18.0)
    aq = 0.07 * np.exp(-(Vp + 20.0) / 20.0)
    bq = 1.0 / (1.0 + np.exp(-(Vp - 10.0) / 10.0))
    return ap / (ap + bp), aq / (aq + bq)


def khh_steady(V, vsh):
    """K_HH1952 (``channel/potassium.py:450-456``)."""
    Vp = V - vsh
    ap = 0.1 / exprel(-(Vp + 10.0) / 10.0)
    bp = 0.125 * np.exp(-(Vp + 20.0) / 80.0)
    return ap / (ap + bp)
