"""Pins the network oracle (``oracle/network.py``) and the network host
logic: the reference network tests' delivery timings restated with their toy
solvers (``network/engine_test.py:109-411``: next-step delivery, multi-step
and non-integer delays with ceil / floor, heterogeneous per-contact delays
and weights, arrivals from several populations summed), the synapse kinetics
against closed forms (ExpSyn decay, the Exp2Syn peak normalised to the
weight, a constant synaptic conductance as a second leak) and the product's
contact tables and delay quantisation against the oracle's."""

import math

import numpy as np
import pytest

from braincell_amd.cell import Cell
from braincell_amd.filter import AllRegion, BranchLocation, RootLocation
from braincell_amd.mech import CableProperty, Channel, Synapse
from braincell_amd.morph import Branch, Morphology
from braincell_amd.network import Connection, Network, delay_steps
from braincell_amd.policy import CVPerBranch
from oracle.network import NetworkOracle, _exp2_factor, _steps

DT = 0.1


def soma_cell(size, V0=-65.0, synapse=None, dend=False, v_th=0.0, g_leak=1e-4, e_leak=-65.0):
    m = Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma"))
    if dend:
        m.attach(0, Branch.from_lengths(lengths=[100.0], radii=[2.0, 1.0], type="basal_dendrite"))
    c = Cell(m, cv_policy=CVPerBranch(1), pop_size=(size,), V_th=v_th)
    c.paint(AllRegion(), CableProperty(resting_potential=V0), Channel("IL", g_max=g_leak, E=e_leak))
    if synapse is not None:
        loc = RootLocation(0.5)
        if dend:
            class _Two:
                def locations(self, morpho):
                    return [(0, 0.5), (1, 0.5)]
            loc = _Two()
        c.place(loc, synapse)
    return c


EXP = Synapse("ExpSyn", tau=2.0, e=0.0, weight=1.0, name="exp")
UP = lambda V: V + 40.0    # spikes on its first step from -10 mV
DOWN = lambda V: V - 1.0   # never spikes


def _toy(pops, cons, n_steps):
    """pops: (name, cell, V0, toy solver); returns g traces of the 'exp' pools."""
    o = NetworkOracle([(n, c.lower(), c.pop_size[0], c.V_th) for n, c, _, _ in pops], cons, DT,
                      solver={n: s for n, _, _, s in pops},
                      quantization=cons[0].get("quantization", "ceil") if cons else "ceil")
    for n, c, V0, _ in pops:
        for cell in o.pops[n]["cells"]:
            cell.V = np.full(cell.C, V0)
    probes = [(n, "exp") for n, c, _, _ in pops if c._synapses]
    _, G, S = o.run(n_steps, pool_probes=probes)
    return G, S


def test_cross_population_delivery_arrives_on_next_step():
    G, S = _toy([("E", soma_cell(2), -10.0, UP), ("I", soma_cell(2, synapse=EXP), -65.0, DOWN)],
                [dict(pre="E", post="I", pre_index=[0], post_index=[1], synapse="exp")], 3)
    g = G[("I", "exp")][:, :, 0]
    assert g.shape == (3, 2)
    assert g[0, 1] == 0.0 and g[1, 1] > 0.0 and g[1, 0] == 0.0
    assert S["E"][0].all() and not S["E"][1:].any()


def test_same_population_recurrent_delivery_uses_next_step():
    G, _ = _toy([("E", soma_cell(2, synapse=EXP), -10.0, UP)],
                [dict(pre="E", post="E", pre_index=[0], post_index=[1], synapse="exp")], 3)
    g = G[("E", "exp")][:, :, 0]
    assert g[0, 1] == 0.0 and g[1, 1] > 0.0


@pytest.mark.parametrize("delay,quant,first", [(0.2, "ceil", 2), (0.15, "ceil", 2), (0.15, "floor", 1)])
def test_delays_and_quantisation(delay, quant, first):
    G, _ = _toy([("E", soma_cell(2), -10.0, UP), ("I", soma_cell(2, synapse=EXP), -65.0, DOWN)],
                [dict(pre="E", post="I", pre_index=[0], post_index=[1], synapse="exp", delay=delay,
                      quantization=quant)], 4)
    g = G[("I", "exp")][:, 1, 0]
    assert np.all(g[:first] == 0.0) and g[first] > 0.0


def test_per_contact_delays_and_weights():
    G, _ = _toy([("E", soma_cell(2), -10.0, UP), ("I", soma_cell(2, synapse=EXP), -65.0, DOWN)],
                [dict(pre="E", post="I", pre_index=[0, 1], post_index=[1, 1], synapse="exp",
                      weight=[0.25, 0.75], delay=[0.0, 0.2])], 4)
    np.testing.assert_allclose(G[("I", "exp")][:, 1, 0], [0.0, 0.25, 1.0, 1.0], atol=1e-12)


def test_arrivals_from_several_populations_sum():
    G, _ = _toy([("A", soma_cell(1), -10.0, UP), ("B", soma_cell(1), -10.0, UP),
                 ("I", soma_cell(2, synapse=EXP), -65.0, DOWN)],
                [dict(pre="A", post="I", pre_index=[0], post_index=[1], synapse="exp", weight=0.25),
                 dict(pre="B", post="I", pre_index=[0], post_index=[1], synapse="exp", weight=0.75)], 2)
    np.testing.assert_allclose(G[("I", "exp")][:, 1, 0], [0.0, 1.0], atol=1e-12)


def test_pool_locations_receive_their_synapse_index():
    G, _ = _toy([("E", soma_cell(1), -10.0, UP), ("I", soma_cell(1, synapse=EXP, dend=True), -65.0, DOWN)],
                [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="exp", synapse_index=[1],
                      weight=0.5)], 2)
    np.testing.assert_allclose(G[("I", "exp")][1, 0], [0.0, 0.5], atol=1e-12)


def test_expsyn_decays_by_exponential_euler():
    dt, tau, w = 0.025, 2.0, 0.3
    pre = soma_cell(1, V0=-10.0)
    post = soma_cell(1, synapse=Synapse("ExpSyn", tau=tau, e=0.0, weight=w, name="exp"))
    o = NetworkOracle([("E", pre.lower(), 1, 0.0), ("I", post.lower(), 1, 0.0)],
                      [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="exp")], dt,
                      solver={"E": UP})
    o.pops["E"]["cells"][0].V[:] = -10.0
    _, G, _ = o.run(40, pool_probes=[("I", "exp")])
    g = G[("I", "exp")][:, 0, 0]
    k = 1.0 - math.expm1(-dt / tau) / (-dt / tau) * dt / tau
    np.testing.assert_allclose(g[1:], w * k ** np.arange(1, 40), rtol=1e-13)
    np.testing.assert_allclose(k, math.exp(-dt / tau), rtol=1e-15)


def test_exp2syn_peak_equals_the_weight():
    for tau1, tau2 in ((0.5, 5.0), (0.2, 10.0), (1.0, 1.0005)):
        t1, f = _exp2_factor(tau1, tau2)
        t = np.linspace(0.0, 20.0 * tau2, 400001)
        g = f * (np.exp(-t / tau2) - np.exp(-t / t1))
        assert abs(g.max() - 1.0) < 1e-6
        tp = (t1 * tau2) / (tau2 - t1) * math.log(tau2 / t1)  # the peak time
        assert abs(f * (math.exp(-tp / tau2) - math.exp(-tp / t1)) - 1.0) < 1e-12
    # the product's kinetic terms carry the same factor
    s = Synapse("Exp2Syn", tau1=0.5, tau2=5.0, weight=1.0)
    assert abs(s.factor() - _exp2_factor(0.5, 5.0)[1]) < 1e-12


def test_constant_synaptic_conductance_acts_as_a_second_leak():
    # tau -> infinity: after the event the synapse is a conductance gs to e
    gs, e, EL, gL = 0.002, 0.0, -65.0, 1e-4
    pre = soma_cell(1, V0=-10.0)
    post = soma_cell(1, synapse=Synapse("ExpSyn", tau=1e12, e=e, weight=gs, name="exp"), g_leak=gL, e_leak=EL)
    o = NetworkOracle([("E", pre.lower(), 1, 0.0), ("I", post.lower(), 1, 100.0)],
                      [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="exp")], 0.025,
                      solver={"E": UP})
    o.pops["E"]["cells"][0].V[:] = -10.0
    V, _, _ = o.run(20000, probe_cvs={"I": [0]})
    area = o.pops["I"]["cells"][0].area_cm2[0]
    G = 1e-3 * gs / area  # mS/cm^2
    Vinf = (gL * 1e3 * EL + G * e) / (gL * 1e3 + G)
    assert abs(V["I"][-1, 0, 0] - Vinf) < 1e-6


def test_product_delay_quantisation_matches_the_oracle():
    d = np.asarray([0.0, 0.05, 0.1, 0.15, 0.1999999999, 0.2, 0.25, 1.0, 3.3])
    for q in ("ceil", "floor"):
        np.testing.assert_array_equal(delay_steps(d, 0.1, q), _steps(d, 0.1, q))
    with pytest.raises(ValueError):
        delay_steps([0.15], 0.1, "strict")
    with pytest.raises(ValueError):
        delay_steps([-1.0], 0.1)


def test_network_api_validation():
    net = Network()
    net.add_population("E", soma_cell(2))
    net.add_population("I", soma_cell(3, synapse=EXP))
    with pytest.raises(ValueError):
        net.add_population("E", soma_cell(1))
    with pytest.raises(ValueError):
        net.add_connection(Connection("E", "X", [0], [0], "exp"))
    with pytest.raises(ValueError):
        net.add_connection(Connection("E", "I", [2], [0], "exp"))
    with pytest.raises(ValueError):
        Connection("E", "I", [0, 1], [0], "exp")
    with pytest.raises(ValueError):
        Cell(Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma")),
             pop_size=(2, 2))
    with pytest.raises(ValueError):
        net.run(dt=0.1, duration=1.0, spike_recording="cells")


def test_product_tables_gather_the_oracle_contacts():
    """The kernel's per-synapse contact lists (CSR by target) hold exactly the
    oracle's delivery blocks' contacts, with their delays and weights."""
    net = Network()
    net.add_population("E", soma_cell(3))
    net.add_population("I", soma_cell(2, synapse=Synapse("Exp2Syn", tau1=0.5, tau2=5.0, weight=0.4, name="exp"),
                                      dend=True))
    net.add_connection(Connection("E", "I", [0, 1, 2, 0], [1, 1, 0, 0], "exp", weight=[0.1, 0.2, 0.3, 0.4],
                                  delay=[0.3, 0.1, 0.25, 1.0], synapse_index=[0, 1, 1, 0]))
    net.add_connection(Connection("I", "I", [0], [1], "exp", delay=0.5))
    tabs = net.build(0.1)
    assert tabs.n_syn == 4 and tabs.n_cells == 5
    got = []
    for key, s in tabs.syn_key.items():
        for c in range(tabs.con_off[s], tabs.con_off[s + 1]):
            got.append((key[0], key[2], int(tabs.con_pre[c]), int(tabs.con_delay[c]), float(tabs.con_w[c])))
    want = [(3 + 1, 0, 0, 3, 0.1), (3 + 1, 1, 1, 1, 0.2), (3 + 0, 1, 2, 3, 0.3), (3 + 0, 0, 0, 10, 0.4),
            (3 + 1, 0, 3, 5, 0.4)]
    assert sorted(got) == sorted(want)
    assert tabs.min_delay == 1 and tabs.max_delay == 10 and tabs.launch_steps == 1
    # synapse rows are grouped per pack and ordered by row
    for p in range(tabs.L.n_packs):
        rows = tabs.syn_row[tabs.syn_pack_off[p]:tabs.syn_pack_off[p + 1]]
        assert np.all(np.diff(rows) > 0)


# ------------------------------------------------------- Markov synapses
def _pool(syn):
    from oracle.network import _Pool

    return _Pool(syn, [0])


def test_ampa_and_gabaa_steps_are_the_exact_solution_for_the_step_drive():
    """g' = alpha D T (1 - g) - beta g is linear in g for the step's drive D,
    so ind_exp_euler is its exact solution over the step:
    g_inf + (g - g_inf) exp(-(alpha D T + beta) dt)."""
    rng = np.random.default_rng(0)
    dt = 0.025
    for syn, (a, b, T) in ((Synapse("AMPA", name="a"), (0.98, 0.18, 0.5)),
                           (Synapse("GABAa", alpha=0.4, T=2.0, name="b"), (0.4, 0.18, 2.0))):
        p = _pool(syn)
        g = 0.0
        for D in rng.choice([0.0, 0.0, 0.0, 1.0, 7.5], 400):
            p.apply(np.asarray([D]))
            p.decay(dt)
            k = a * D * T + b
            ginf = a * D * T / k
            g = ginf + (g - ginf) * math.exp(-k * dt)
            assert abs(p.g[0] - g) <= 1e-14
        assert 0.0 < g < 1.0


def test_nmda_steps_every_state_from_the_step_start_values():
    """ind_exp_euler takes each state's derivative and diagonal Jacobian at
    the step-start values: g steps as the exact solution with x frozen at
    its old value (not the new one), x as the exact solution for the
    drive; and over many steps the scheme converges to the ODE at first
    order in dt."""
    from scipy.integrate import solve_ivp

    syn = Synapse("NMDA", alpha1=2.0, beta1=0.3, alpha2=1.0, beta2=0.5, T=1.0, name="n")
    p = _pool(syn)
    p.A[:] = 0.6  # x
    p.B[:] = 0.3  # g
    D, dt = 2.0, 0.05
    p.apply(np.asarray([D]))
    p.decay(dt)
    k = 2.0 * 0.6 + 0.3
    g_old_x = 2.0 * 0.6 / k + (0.3 - 2.0 * 0.6 / k) * math.exp(-k * dt)
    kx = 1.0 * D + 0.5
    x1 = D / kx + (0.6 - D / kx) * math.exp(-kx * dt)
    assert abs(p.g[0] - g_old_x) < 1e-14 and abs(p.A[0] - x1) < 1e-14
    k1 = 2.0 * x1 + 0.3
    assert abs(p.g[0] - (2.0 * x1 / k1 + (0.3 - 2.0 * x1 / k1) * math.exp(-k1 * dt))) > 1e-4

    def rhs(t, y):
        g, x = y
        return [2.0 * x * (1 - g) - 0.3 * g, 1.0 * D * (1 - x) - 0.5 * x]

    ref = solve_ivp(rhs, (0.0, 5.0), [0.0, 0.0], rtol=1e-12, atol=1e-14).y[:, -1]
    errs = []
    for dt in (0.05, 0.025, 0.0125):
        q = _pool(syn)
        for _ in range(int(round(5.0 / dt))):
            q.apply(np.asarray([D]))
            q.decay(dt)
        errs.append(abs(q.g[0] - ref[0]))
    assert errs[0] > errs[1] > errs[2] and 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3


def test_markov_synapse_current_is_a_density_at_its_cv():
    """With beta = 0 an AMPA synapse keeps the conductance one event gives
    it: g_max g (mS/cm^2, no area division) then acts as a second leak."""
    EL, gL = -65.0, 1e-4
    pre = soma_cell(1, V0=-10.0)
    post = soma_cell(1, synapse=Synapse("AMPA", alpha=1.0, beta=0.0, T=1.0, g_max=0.05, E_rev=0.0, name="a"),
                     g_leak=gL, e_leak=EL)
    dt = 0.025
    o = NetworkOracle([("E", pre.lower(), 1, 0.0), ("I", post.lower(), 1, 100.0)],
                      [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="a", weight=4.0)], dt,
                      solver={"E": UP})
    o.pops["E"]["cells"][0].V[:] = -10.0
    V, G, _ = o.run(20000, probe_cvs={"I": [0]}, pool_probes=[("I", "a")])
    g = G[("I", "a")][-1, 0, 0]
    assert abs(g - (1.0 - math.exp(-4.0 * dt))) < 1e-14  # one event: D = 4 for one step
    Gs = 0.05 * g
    Vinf = (gL * 1e3 * EL + Gs * 0.0) / (gL * 1e3 + Gs)
    assert abs(V["I"][-1, 0, 0] - Vinf) < 1e-6


def test_markov_synapses_need_a_connection_weight_and_lower_to_kernel_terms():
    from braincell_amd.mech import SYN_MARKOV1, SYN_MARKOV2

    pre = soma_cell(1, V0=-10.0)
    post = soma_cell(1, synapse=Synapse("AMPA", name="a"))
    with pytest.raises(ValueError, match="weight"):
        NetworkOracle([("E", pre.lower(), 1, 0.0), ("I", post.lower(), 1, 0.0)],
                      [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="a")], DT)
    net = Network()
    net.add_population("E", soma_cell(1))
    net.add_population("I", soma_cell(1, synapse=Synapse("NMDA", name="n")))
    net.add_connection(Connection("E", "I", [0], [0], "n"))
    with pytest.raises(ValueError, match="weight"):
        net.build(0.025)
    with pytest.raises(ValueError):
        Synapse("AMPA", weight=1.0)
    a = Synapse("AMPA", g_max=0.4, T=2.0).kernel_par(0.025, 1.5, 1e-5)
    assert a == pytest.approx((0.025 / 1.5 * 0.4, 0.0, SYN_MARKOV1, 0.98 * 2.0, 0.18, 0.0, 0.0, 0.0))
    n = Synapse("NMDA", T=0.5).kernel_par(0.025, 1.0, 1e-5)
    assert n == pytest.approx((0.025, 0.0, SYN_MARKOV2, 2.0, 0.01, 0.5, 0.5, 0.0))
    gab = Synapse("GABAa").resolved()
    assert gab["E_rev"] == -70.0 and gab["alpha"] == 0.53 and gab["beta"] == 0.18


def test_a_new_run_drops_events_in_flight_and_keeps_the_state():
    """``Network.run`` resets the delivery queues before every run
    (``engine.py:602-609``): a spike sent at the last step of one run with a
    two-step delay never arrives in the next, while cell states carry on."""
    from oracle.network import NetworkOracle

    def make():
        pre, post = soma_cell(1), soma_cell(1, synapse=EXP)
        o = NetworkOracle([("E", pre.lower(), 1, 0.0), ("I", post.lower(), 1, 0.0)],
                          [dict(pre="E", post="I", pre_index=[0], post_index=[0], synapse="exp", delay=0.2)], DT,
                          solver={"E": UP, "I": DOWN})
        for c in o.pops["E"]["cells"]:
            c.V = np.full(c.C, -10.0)
        for c in o.pops["I"]["cells"]:
            c.V = np.full(c.C, -65.0)
        return o

    whole = make()
    _, G, _ = whole.run(4, pool_probes=[("I", "exp")])
    assert G[("I", "exp")][2, 0, 0] == pytest.approx(1.0)
    split = make()
    split.run(1, pool_probes=[("I", "exp")])
    split.reset_delivery()
    V2, G2, _ = split.run(3, probe_cvs={"I": [0]}, pool_probes=[("I", "exp")], step0=1)
    assert np.all(G2[("I", "exp")] == 0.0)
    np.testing.assert_allclose(V2["I"][:, 0, 0], [-67.0, -68.0, -69.0])  # the toy solver kept stepping
