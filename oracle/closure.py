"""FP64 restatement of the learned gating closure (test infrastructure).

The closure has no counterpart in the reference.  This module restates, in
NumPy FP64, what the closure kernels compute, so the GPU kernels can be
checked against it: the plain exponential-Euler gate update that the
training targets subtract (``quad/_exp_euler.py:330-478`` for the update,
the channel files ``sodium.py:134-170``, ``potassium.py:189-214``,
``calcium.py:880-920`` for the rates), the feature vector, and the MLP
8 -> 64 -> 64 -> 5 (tanh) with its increment-weighted MSE loss
sum_o (delta_o c_o - y_o)^2 and gradient by backpropagation.
"""

from __future__ import annotations

import numpy as np

from . import channels as ch

NIN, NH, NOUT = 8, 64, 5
SIZES = ((NIN, NH), (NH, NH), (NH, NOUT))


def unpack(P):
    """Flat parameters -> [(W, b)] with W[in, out]."""
    P = np.asarray(P, dtype=np.float64)
    out, o = [], 0
    for fi, fo in SIZES:
        W = P[o:o + fi * fo].reshape(fi, fo)
        o += fi * fo
        b = P[o:o + fo]
        o += fo
        out.append((W, b))
    assert o == P.size
    return out


def forward(P, Xn):
    """Network output for normalised features Xn [n, 8]; also the hidden
    activations."""
    (W1, b1), (W2, b2), (W3, b3) = unpack(P)
    h1 = np.tanh(Xn @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ W3 + b3, h1, h2


def loss_and_grad(P, Xn, Dn, Yn):
    """mean_b sum_o (Dn * net - Yn)^2 and its gradient (flat, FP64)."""
    (W1, b1), (W2, b2), (W3, b3) = unpack(P)
    y, h1, h2 = forward(P, Xn)
    B = Xn.shape[0]
    d = Dn * y - Yn
    loss = float(np.sum(d * d) / B)
    d3 = 2.0 * d * Dn / B
    gW3, gb3 = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ W3.T) * (1.0 - h2 * h2)
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * (1.0 - h1 * h1)
    gW1, gb1 = Xn.T @ d1, d1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def plain_gates(V, x, dt, na_vsh=-50.0, kdr_vsh=-50.0, cal_vsh=0.0, na_phi=1.0, kdr_phi=1.0, cal_phi_p=1.0,
                cal_phi_q=1.0):
    """Exponential-Euler update of (Na p, Na q, KDR p, CaL p, CaL q) at V."""
    ap, bp, aq, bq = ch.na_ba2002(V, na_vsh)
    kp, kb = ch.kdr_ba2002(V, kdr_vsh)
    pi, pt, qi, qt = ch.cal_is2008(V, cal_vsh)
    return np.stack([ch.ab_update(x[0], ap, bp, na_phi, dt), ch.ab_update(x[1], aq, bq, na_phi, dt),
                     ch.ab_update(x[2], kp, kb, kdr_phi, dt), ch.it_update(x[3], pi, pt, cal_phi_p, dt),
                     ch.it_update(x[4], qi, qt, cal_phi_q, dt)])


def features(V1, x0, Ci1, dt):
    """Raw feature rows (V_{n+1}, 5 gates at the step start, ln Ci_{n+1}, dt)."""
    n = np.asarray(V1).shape[0]
    return np.column_stack([V1, np.asarray(x0).T, np.log(Ci1), np.full(n, dt)])
