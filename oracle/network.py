"""Oracle restatement (FP64) of a network run: populations of cells,
ExpSyn / Exp2Syn / AMPA / GABAa / NMDA synapses and delayed spike delivery
through per-delay ring buffers.  Test infrastructure only.

Per step n (``network/engine.py:474-548``):

1. ``write_arrivals`` (``delivery.py:317-359``): every delivery block (the
   contacts of one connection with one delay d) reads its ring slot at the
   cursor, sums it into the target pool's ``pre_spike`` and clears it;
2. ``_begin_step`` -> ``apply_discrete_events`` (``synapse/markov.py:
   100-102, 180-184``): ExpSyn g += drive, Exp2Syn A, B += drive * factor;
   the Markov types keep the drive for their derivative (``pre_drive``);
3. the cell step (``StaggeredOracle.step``) with each synapse's current
   g (V - e) / area at its CV (``_multi_compartment/currents.py:138-163``),
   then the synapse states by ind_exp_euler (``cell.py:2506-2527``):
   y + dt exprel(-dt / tau) (-y / tau);
4. ``get_spike`` (``_base.py:421-431``) per CV, reduced per cell with any
   (``delivery.py:103-123``): V_n < V_th <= V_{n+1};
5. ``enqueue_future_events`` (``delivery.py:362-397``): each block adds
   w * spike(pre) at ring slot (cursor + d) mod (d + 1); cursors advance.

External drive (``_multi_compartment/cell.py:2426-2487``): the drive a
synapse takes at step n is the arrivals plus the NetStims placed at its point
(``_compute/layouts.py:906-941``: with local = t - start and
k = round(local / interval), weight when local >= 0, |local - k interval| <=
1e-9 and 0 <= k < number) plus every bound input source() * weight (after
transform).  ``drive_mode`` says at which time t they are evaluated:
"network" t_n (``network/engine.py:491-504``), "run" t_{max(n-1, 0)}
(``Cell.run``, ``_multi_compartment/run.py:108-134``: one preparation
before the loop at the start time, then one after each step at that step's
time) and "update" t_{n-1} (``Cell.update``, ``cell.py:2298-2329``; none at
the first step).

Delays are ``max(ceil(delay / dt - 1e-12), 1)`` steps after snapping to the
nearest integer within 1e-7 (``lowering.py:185-228``).  ``solver`` (a callable
or a dict population -> callable) replaces step 3 with a toy voltage map, as
the reference's network tests do; their synapses then never decay.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import exprel

from .staggered import StaggeredOracle


def _steps(delay_ms, dt, quantization="ceil"):
    raw = np.asarray(delay_ms, dtype=float) / dt
    r = np.rint(raw)
    raw = np.where(np.isclose(raw, r, rtol=1e-7, atol=1e-7), r, raw)
    if quantization == "ceil":
        s = np.ceil(raw - 1e-12)
    elif quantization == "floor":
        s = np.floor(raw + 1e-12)
    else:
        s = np.rint(raw)
    return np.maximum(s.astype(np.int64), 1)


def _exp2_factor(tau1, tau2):
    """markov.py:152-162."""
    ratio = tau1 / tau2
    t1 = tau2 * 0.9999 if ratio > 0.9999 else (tau2 * 1e-9 if ratio < 1e-9 else tau1)
    tp = (t1 * tau2) / (tau2 - t1) * math.log(tau2 / t1)
    return t1, 1.0 / (-math.exp(-(tp / t1)) + math.exp(-(tp / tau2)))


def _ind_exp_euler(y, f, lam, dt):
    """``quad/_exp_euler.py:455-463``: y + dt exprel(dt lam) f."""
    return y + dt * exprel(dt * lam) * f


class _Pool:
    """One synapse pool of one cell: states per location.

    ExpSyn / Exp2Syn (``markov.py:65-200``): A, B in uS, g = B - A, events
    jump A and B (``apply_discrete_events``), the states then decay.
    AMPA / GABAa (``markov.py:203-307``): g' = alpha D T (1 - g) - beta g;
    NMDA (``markov.py:310-360``): g' = alpha1 x (1 - g) - beta1 g,
    x' = alpha2 D T (1 - x) - beta2 x; D is the step's summed weighted events
    (``pre_drive``), used only by ``compute_derivative`` after the voltage
    step; every state steps from the step-start values (ind_exp_euler:
    each derivative and its own diagonal Jacobian at the old states,
    ``_exp_euler.py:427-470``).  Their current is a density
    g_max g (E_rev - V) (mS/cm^2 * mV; ``currents.py:138-163`` adds it at
    the point without an area division)."""

    def __init__(self, syn, cvs):
        p = syn.resolved()
        self.kind = syn.synapse_type
        self.markov = self.kind in ("AMPA", "GABAa", "NMDA")
        self.p = {k: float(v) for k, v in p.items()}
        self.e = float(p["E_rev"] if self.markov else p["e"])
        self.weight = float(p["weight"]) if "weight" in p else None
        self.cvs = np.asarray(cvs, dtype=np.int64)
        n = len(cvs)
        self.A = np.zeros(n)  # Exp2Syn A; NMDA x
        self.B = np.zeros(n)  # ExpSyn / Exp2Syn B; Markov g
        self.D = np.zeros(n)
        if self.kind == "ExpSyn":
            self.tau_a, self.tau_b, self.fa, self.fb = 1.0, float(p["tau"]), 0.0, 1.0
        elif self.kind == "Exp2Syn":
            t1, f = _exp2_factor(float(p["tau1"]), float(p["tau2"]))
            self.tau_a, self.tau_b, self.fa, self.fb = t1, float(p["tau2"]), f, f

    @property
    def g(self):
        return self.B if self.markov else self.B - self.A

    def conductance_density(self, area_cm2):
        """mS/cm^2 at each location: g_max g, or g (uS) / area."""
        if self.markov:
            return self.p["g_max"] * self.B
        return 1e-3 * self.g / area_cm2

    def apply(self, drive):
        if self.markov:
            self.D = np.asarray(drive, dtype=float)
            return
        self.A = self.A + drive * self.fa
        self.B = self.B + drive * self.fb

    def decay(self, dt):
        if self.kind in ("AMPA", "GABAa"):
            a, b, T = self.p["alpha"], self.p["beta"], self.p["T"]
            g = self.B
            self.B = _ind_exp_euler(g, a * self.D * T * (1 - g) - b * g, -(a * self.D * T + b), dt)
        elif self.kind == "NMDA":
            a1, b1, a2, b2, T = (self.p[k] for k in ("alpha1", "beta1", "alpha2", "beta2", "T"))
            g, x = self.B, self.A
            self.B = _ind_exp_euler(g, a1 * x * (1 - g) - b1 * g, -(a1 * x + b1), dt)
            self.A = _ind_exp_euler(x, a2 * self.D * T * (1 - x) - b2 * x, -(a2 * self.D * T + b2), dt)
        else:
            self.A = self.A + dt * exprel(-dt / self.tau_a) * (-self.A / self.tau_a)
            self.B = self.B + dt * exprel(-dt / self.tau_b) * (-self.B / self.tau_b)
        self.D = np.zeros_like(self.D)


class NetworkOracle:
    """``populations``: list of (name, CellModel, size, V_th); ``connections``:
    list of dicts with pre, post, pre_index, post_index, synapse,
    synapse_index, weight (uS or None) and delay (ms)."""

    def __init__(self, populations, connections, dt, *, quantization="ceil", solver=None, V_init=None,
                 drive_mode="network", bindings=None):
        self.dt = float(dt)
        self.solver = solver
        if drive_mode not in ("network", "run", "update"):
            raise ValueError(drive_mode)
        self.drive_mode = drive_mode
        self.bindings = bindings or {}  # population -> {pool: [(source, weight, transform)]}
        self.pops = {}
        self.order = []
        for name, model, size, v_th in populations:
            cells, pools = [], []
            for _ in range(size):
                c = StaggeredOracle(model)
                c.reset_state(model.V0 if V_init is None else V_init)
                cp = {}
                for (pname, syn, sites) in model.synapses:
                    cvs = []
                    for (b, x) in sites:
                        kind, idx = c._locate(b, x)
                        if kind != "mid":
                            raise ValueError("synapses sit inside a CV")
                        cvs.append(idx)
                    cp[pname] = _Pool(syn, cvs)
                cells.append(c)
                pools.append(cp)
            stims = []  # (start, number, interval, weight, CV)
            for (ns, sites) in getattr(model, "netstims", []) or []:
                for (b, x) in sites:
                    kind, idx = cells[0]._locate(b, x)
                    if kind == "mid":
                        stims.append((ns.start, ns.number, ns.interval, ns.weight, idx))
            self.pops[name] = {"cells": cells, "pools": pools, "v_th": float(v_th), "size": size, "stims": stims}
            self.order.append(name)
        # delivery blocks: (connection, delay) groups with a ring [d + 1, post, n_active]
        self.blocks = []
        for con in connections:
            post = self.pops[con["post"]]
            n_active = len(post["pools"][0][con["synapse"]].cvs)
            n = len(con["pre_index"])
            w = con.get("weight")
            if w is None:
                w0 = post["pools"][0][con["synapse"]].weight
                if w0 is None:  # lowering.py:130-133
                    raise ValueError("Connection weight is None, but the target synapse has no default 'weight'.")
                w = np.full(n, w0)
            w = np.broadcast_to(np.asarray(w, dtype=float), (n,))
            d = _steps(np.broadcast_to(np.asarray(con.get("delay", 0.0), dtype=float), (n,)), self.dt, quantization)
            sidx = np.asarray(con.get("synapse_index", np.zeros(n, dtype=np.int64)), dtype=np.int64)
            for dd in np.unique(d):
                m = d == dd
                self.blocks.append({
                    "pre": con["pre"], "post": con["post"], "synapse": con["synapse"], "d": int(dd),
                    "pre_index": np.asarray(con["pre_index"])[m], "post_index": np.asarray(con["post_index"])[m],
                    "syn_index": sidx[m], "w": w[m],
                    "ring": np.zeros((int(dd) + 1, post["size"], n_active)), "cursor": 0})
        self.pre_spike = {}
        self.spikes = {name: np.zeros(self.pops[name]["size"], dtype=bool) for name in self.order}
        self.cv_spikes = {name: np.zeros((self.pops[name]["size"], self.pops[name]["cells"][0].C), dtype=bool)
                          for name in self.order}

    @staticmethod
    def _netstim(start, number, interval, weight, t):
        local = t - start
        k = np.round(local / interval)
        fired = local >= 0.0 and abs(local - k * interval) <= 1e-9 and 0 <= k < number
        return weight if fired else 0.0

    def _external(self, name, n):
        """{pool: [size, n_loc]} NetStim + bound drive of step n."""
        P = self.pops[name]
        src = n if self.drive_mode == "network" else (max(n - 1, 0) if self.drive_mode == "run" else n - 1)
        out = {}
        if src < 0:
            return out
        t = src * self.dt
        for pname, pool in P["pools"][0].items():
            d = np.zeros((P["size"], len(pool.cvs)))
            for (start, number, interval, w, cv) in P["stims"]:
                amp = self._netstim(start, number, interval, w, t)
                for j, pcv in enumerate(pool.cvs):
                    if pcv == cv:
                        d[:, j] += amp
            for source, weight, transform in self.bindings.get(name, {}).get(pname, []):
                v = source() if callable(source) else source
                if transform is not None:
                    v = transform(v)
                d += np.broadcast_to(np.asarray(v, dtype=float) * np.asarray(weight, dtype=float), d.shape)
            out[pname] = d
        return out

    def step(self, n):
        t = n * self.dt
        # 1. due arrivals
        grouped = {}
        for b in self.blocks:
            key = (b["post"], b["synapse"])
            arr = b["ring"][b["cursor"]].copy()
            grouped[key] = arr if key not in grouped else grouped[key] + arr
            b["ring"][b["cursor"]] = 0.0
        for key, arr in grouped.items():
            self.pre_spike[key] = arr
        for name in self.order:
            P = self.pops[name]
            ext = self._external(name, n)
            for i in range(P["size"]):
                c, pools = P["cells"][i], P["pools"][i]
                # 2. discrete events: arrivals, then NetStim and bound drive
                for pname, pool in pools.items():
                    drive = self.pre_spike.get((name, pname))
                    total = np.zeros(len(pool.cvs)) if drive is None else drive[i].copy()
                    if pname in ext:
                        total = total + ext[pname][i]
                    if drive is not None or pname in ext:
                        pool.apply(total)
                V_old = c.V.copy()
                # 3. cell step with the synapses' conductance, then their decay
                toy = self.solver.get(name) if isinstance(self.solver, dict) else self.solver
                if toy is not None:
                    c.V = toy(c.V)
                else:
                    Gs = np.zeros(c.C)
                    GEs = np.zeros(c.C)
                    for pool in pools.values():
                        gk = pool.conductance_density(c.area_cm2[pool.cvs])
                        np.add.at(Gs, pool.cvs, gk)
                        np.add.at(GEs, pool.cvs, gk * pool.e)
                    c.point_conductance = (Gs, GEs)
                    c.step(t, self.dt)
                    for pool in pools.values():
                        pool.decay(self.dt)
                # 4. spike: any CV crossing V_th upwards (get_spike per CV, reduced per cell
                #    for the "population" record; the "full" record keeps every CV)
                cross = (V_old < P["v_th"]) & (c.V >= P["v_th"])
                self.cv_spikes[name][i] = cross
                self.spikes[name][i] = bool(np.any(cross))
        # 5. enqueue and advance
        for b in self.blocks:
            spk = self.spikes[b["pre"]][b["pre_index"]].astype(float)
            ev = np.zeros(b["ring"].shape[1:])
            np.add.at(ev, (b["post_index"], b["syn_index"]), b["w"] * spk)
            tgt = (b["cursor"] + b["d"]) % b["ring"].shape[0]
            b["ring"][tgt] += ev
        for b in self.blocks:
            b["cursor"] = (b["cursor"] + 1) % b["ring"].shape[0]

    def reset_delivery(self):
        """Empty every delivery ring and rewind its cursor, as
        ``Network.run`` does before each run (``engine.py:602-609``): events
        still in flight are dropped; cell and synapse states are kept."""
        for b in self.blocks:
            b["ring"][:] = 0.0
            b["cursor"] = 0
        self.pre_spike = {}

    def run(self, n_steps, probe_cvs=None, pool_probes=None, record_cv_spikes=False, step0=0):
        """Returns (V traces {pop: [n_steps, size, n_cv_probe]}, pool traces
        {(pop, pool[, field]): [n_steps, size, n_active]}, spikes {pop:
        [n_steps, size]}).  A pool probe (pop, pool) records g; (pop, pool,
        field) records "g", "A" or "B" (Exp2Syn), "x" (NMDA, = A).  With
        ``record_cv_spikes`` the per-CV crossings land in
        ``self.cv_spike_trace`` {pop: [n_steps, size, C]}.  ``step0``: the
        global index of the first step (the clock of a continued run).
        Probes, pool traces and spikes are indexed from the run's first step."""
        probe_cvs = probe_cvs or {}
        pool_probes = pool_probes or []
        V = {k: np.zeros((n_steps, self.pops[k]["size"], len(v))) for k, v in probe_cvs.items()}
        G = {k: np.zeros((n_steps, self.pops[k[0]]["size"], len(self.pops[k[0]]["pools"][0][k[1]].cvs)))
             for k in pool_probes}
        S = {k: np.zeros((n_steps, self.pops[k]["size"]), dtype=bool) for k in self.order}
        if record_cv_spikes:
            self.cv_spike_trace = {k: np.zeros((n_steps,) + self.cv_spikes[k].shape, dtype=bool) for k in self.order}
        for n in range(n_steps):
            self.step(step0 + n)
            for k, cvs in probe_cvs.items():
                for i, c in enumerate(self.pops[k]["cells"]):
                    V[k][n, i] = c.V[cvs]
            for key in pool_probes:
                pk, pn = key[0], key[1]
                field = key[2] if len(key) > 2 else "g"
                for i, pools in enumerate(self.pops[pk]["pools"]):
                    pool = pools[pn]
                    G[key][n, i] = pool.g if field == "g" else (pool.A if field in ("A", "x") else pool.B)
            for k in self.order:
                S[k][n] = self.spikes[k]
                if record_cv_spikes:
                    self.cv_spike_trace[k][n] = self.cv_spikes[k]
        return V, G, S
