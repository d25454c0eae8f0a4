"""Oracle restatement (FP64) of ``backward_euler`` on a passive
multi-compartment cell.  Test infrastructure only.

``backward_euler_step`` (``quad/_backward_euler.py:71-160``) runs
``apply_standard_solver_step(_backward_euler, cell, t, dt, merging='stack')``
(``quad/_util.py:196-262``): the cell's differential states are stacked
along a new last axis (``_util.py:116-120``), which for a passive cell is V
alone, y = V[..., None] with one value per CV.  The vector field is
``Cell.compute_derivative`` (``cell.py:1892-1999``):

    f(V) = total_membrane_current(V, t) / Cm - V K^T     (K: build_cv_axial_operator,
                                                          _staggered.py:393-413)

with the membrane current of the leak channels (g (E - V), IL's own E,
K_Leak at E_K) plus the midpoint clamps (``currents.py:166-193``).
``_backward_euler`` (``_backward_euler.py:31-67``) takes
A = jacrev_last_dim(f)(y) (``_util.py:265-310``: one VJP per last-axis
component with the cotangent 1 at every position, so for M = 1 the entry
at position i is sum_q df_q / dV_i) and solves (1 - dt A_i) dy_i =
dt f_i(y) per position; then ``post_integral`` adds the delta inputs (none).
The time passed to f is the step's start time.
"""

from __future__ import annotations

import numpy as np

from .discretize import schur_cv_operator
from .staggered import StaggeredOracle

PASSIVE = ("IL", "K_Leak")


class CableBEOracle:
    def __init__(self, model):
        extra = set(model.channels) - set(PASSIVE)
        if extra or model.ca is not None or model.synapses:
            raise ValueError("the stacked-state backward Euler needs V to be the only state")
        self.cell = StaggeredOracle(model)
        self.K = schur_cv_operator(self.cell.st)  # [C, C] ms^-1
        c = self.cell
        self.C = c.C
        self.cm = c.cm
        self.gE = []  # (g mS/cm^2 per CV, E)
        for name, g in c.g.items():
            E = c.params[name].get("E", -70.0) if name == "IL" else c.E_K
            self.gE.append((g, float(E)))
        self.V = None

    def reset_state(self, V0):
        self.V = np.broadcast_to(np.asarray(V0, dtype=np.float64), (self.C,)).copy()

    def f(self, V, t):
        I = self.cell.clamp_density(t)
        for g, E in self.gE:
            I = I + g * (E - V)
        return I / self.cm - self.K @ V

    def jacobian_colsum(self):
        """sum over output positions q of df_q / dV_i (the all-ones VJP)."""
        d_mem = np.zeros(self.C)
        for g, _ in self.gE:
            d_mem = d_mem - g / self.cm
        return d_mem - self.K.sum(axis=0)

    def step(self, t, dt):
        J = self.jacobian_colsum()
        self.V = self.V + dt * self.f(self.V, t) / (1.0 - dt * J)

    def run(self, dt, n_steps, probe_cvs=None, step0=0):
        probe_cvs = list(range(self.C)) if probe_cvs is None else list(probe_cvs)
        out = np.empty((n_steps, len(probe_cvs)))
        for n in range(n_steps):
            self.step((step0 + n) * dt, dt)
            out[n] = self.V[probe_cvs]
        return out
