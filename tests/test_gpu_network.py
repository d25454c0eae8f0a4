"""GPU parity of network runs (synapses and delayed spike delivery on the
staggered kernel) against the network oracle."""

import numpy as np
import pytest

from tests.helpers import rmse

pytestmark = pytest.mark.gpu


def _cell(size, amp, *, dend=False, syn=None, precision=64):
    from braincell_amd.cell import Cell
    from braincell_amd.filter import AllRegion, RootLocation
    from braincell_amd.mech import CableProperty, Channel, CurrentClamp, StateProbe
    from braincell_amd.morph import Branch, Morphology
    from braincell_amd.policy import CVPerBranch

    m = Morphology.from_root(Branch.from_lengths(lengths=[20.0], radii=[10.0, 10.0], type="soma"))
    if dend:
        m.attach(0, Branch.from_lengths(lengths=[150.0], radii=[1.5, 0.8], type="basal_dendrite"))
    c = Cell(m, cv_policy=CVPerBranch(3), pop_size=(size,), precision=precision)
    c.paint(AllRegion(), CableProperty(resting_potential=-65.0, axial_resistivity=150.0),
            Channel("IL", g_max=3e-4, E=-54.4), Channel("Na_HH1952", g_max=0.12), Channel("K_HH1952", g_max=0.036))
    if amp:
        c.place(RootLocation(0.5), CurrentClamp(delay=0.5, durations=[30.0], amplitudes=[amp]))
    if syn is not None:
        class _Sites:
            def locations(self, morpho):
                return [(0, 0.5), (1, 0.5)] if dend else [(0, 0.5)]
        c.place(_Sites(), syn)
    c.place(RootLocation(0.5), StateProbe("v"))
    return c


def _net(precision):
    from braincell_amd.mech import Synapse
    from braincell_amd.network import Connection, Network

    net = Network()
    net.add_population("E1", _cell(3, 0.25, precision=precision))
    net.add_population("E2", _cell(2, 0.6, precision=precision))
    net.add_population("I", _cell(4, 0.0, dend=True, precision=precision,
                                  syn=Synapse("Exp2Syn", tau1=0.5, tau2=4.0, e=0.0, weight=0.01, name="ampa")))
    net.add_population("J", _cell(2, 0.02, precision=precision,
                                  syn=Synapse("ExpSyn", tau=3.0, e=-80.0, weight=0.02, name="gaba")))
    net.add_connection(Connection("E1", "I", [0, 1, 2, 0, 2], [0, 1, 2, 3, 3], "ampa",
                                  weight=[0.004, 0.006, 0.008, 0.01, 0.003], delay=[0.5, 1.0, 0.25, 2.0, 0.8],
                                  synapse_index=[0, 1, 0, 1, 1]))
    net.add_connection(Connection("E2", "I", [0, 1], [0, 3], "ampa", weight=0.02, delay=0.3))
    net.add_connection(Connection("I", "J", [0, 1, 2, 3], [0, 0, 1, 1], "gaba", delay=1.0))
    net.add_connection(Connection("I", "I", [1], [2], "ampa", weight=0.015, delay=0.6, synapse_index=[1]))
    return net


def _oracle(net, dt, n_steps, pool_probes=None):
    from oracle.network import NetworkOracle

    pops = [(n, p.cell.lower(), p.size, p.cell.V_th) for n, p in net.populations.items()]
    cons = [dict(pre=c.pre_population, post=c.post_population, pre_index=c.pre_index, post_index=c.post_index,
                 synapse=c.synapse, synapse_index=c.synapse_index, weight=c.weight, delay=c.delay)
            for c in net.all_connections()]
    o = NetworkOracle(pops, cons, dt)
    cvs = {n: [o.pops[n]["cells"][0]._locate(0, 0.5)[1]] for n in net.populations}
    return o.run(n_steps, probe_cvs=cvs, pool_probes=pool_probes)


def test_network_fp64_matches_oracle(cuda):
    dt, T = 0.025, 15.0
    n = int(round(T / dt))
    print(f"network fp64, {n} steps", flush=True)
    net = _net(64)
    res = net.run(dt=dt, duration=T, spike_recording="population")
    V, _, S = _oracle(net, dt, n)
    nsp = 0
    for name in net.populations:
        got, ref = res.traces[name]["v"], V[name][:, :, 0]
        err = np.abs(got - ref).max()
        print(name, "max |dV|", err, "spikes", int(S[name].sum()))
        assert err < 1e-7
        assert np.array_equal(res.spikes[name], S[name])
        nsp += int(S[name].sum())
    assert S["E1"].any() and S["E2"].any() and S["I"].any()  # events travel E -> I -> J
    ref_j = V["J"][:, :, 0]
    assert np.ptp(ref_j) > 1.0 and nsp > 5


def test_network_fp32_rmse_and_spike_steps(cuda):
    dt, T = 0.025, 15.0
    n = int(round(T / dt))
    print(f"network fp32, {n} steps", flush=True)
    net = _net(32)
    res = net.run(dt=dt, duration=T, spike_recording="population")
    V, _, S = _oracle(net, dt, n)
    for name in net.populations:
        got, ref = res.traces[name]["v"], V[name][:, :, 0]
        worst = max(rmse(got[:, i], ref[:, i]) for i in range(ref.shape[1]))
        print(name, "fp32 worst RMSE", worst)
        assert worst <= 0.05
        for i in range(ref.shape[1]):
            a, b = np.flatnonzero(res.spikes[name][:, i]), np.flatnonzero(S[name][:, i])
            assert len(a) == len(b) and (len(a) == 0 or np.max(np.abs(a - b)) <= 1)


def test_network_launches_are_bounded_by_the_shortest_delay(cuda):
    """The shortest delay, 0.25 ms at dt 0.025, allows 10-step launches; the
    same network with one-step launches gives identical results."""
    from braincell_amd.network import Network

    net = _net(64)
    tabs = net.build(0.025)
    assert tabs.min_delay == 10 and tabs.launch_steps == 10
    a = net.run(dt=0.025, duration=5.0, spike_recording="population")
    net.reset_state()
    b = net.run(dt=0.025, duration=5.0, chunk=1, spike_recording="population")
    for name in net.populations:
        assert np.array_equal(a.traces[name]["v"], b.traces[name]["v"])
        assert np.array_equal(a.spikes[name], b.spikes[name])
    assert isinstance(net, Network)


def _markov_net(precision):
    """AMPA and NMDA pools on two sites of the I cells, GABAa on J
    (``synapse/markov.py:203-360``); weights are dimensionless drives."""
    from braincell_amd.mech import StateProbe, Synapse
    from braincell_amd.network import Connection, Network

    class _Sites:
        def locations(self, morpho):
            return [(0, 0.5), (1, 0.5)]

    i_cell = _cell(4, 0.0, dend=True, precision=precision)
    i_cell.place(_Sites(), Synapse("AMPA", g_max=3.0, name="ampa"))
    i_cell.place(_Sites(), Synapse("NMDA", g_max=0.3, beta1=0.05, name="nmda"))
    j_cell = _cell(2, 0.02, precision=precision, syn=Synapse("GABAa", g_max=2.0, name="gaba"))
    net = Network()
    net.add_population("E1", _cell(3, 0.25, precision=precision))
    net.add_population("E2", _cell(2, 0.6, precision=precision))
    net.add_population("I", i_cell)
    net.add_population("J", j_cell)
    net.add_connection(Connection("E1", "I", [0, 1, 2, 0, 2], [0, 1, 2, 3, 3], "ampa",
                                  weight=[40.0, 60.0, 80.0, 50.0, 30.0], delay=[0.5, 1.0, 0.25, 2.0, 0.8],
                                  synapse_index=[0, 1, 0, 1, 1]))
    net.add_connection(Connection("E2", "I", [0, 1], [0, 3], "nmda", weight=20.0, delay=0.3,
                                  synapse_index=[1, 0]))
    net.add_connection(Connection("I", "J", [0, 1, 2, 3], [0, 0, 1, 1], "gaba", weight=10.0, delay=1.0))
    net.add_connection(Connection("I", "I", [1], [2], "ampa", weight=5.0, delay=0.6, synapse_index=[1]))
    return net


@pytest.mark.parametrize("precision", [64, 32])
def test_network_with_ampa_nmda_gabaa_matches_oracle(cuda, precision):
    """FP64: every probe within 1e-7 mV of the oracle and identical spike
    steps; FP32: soma RMSE <= 0.05 mV and spike steps within one step."""
    dt, T = 0.025, 15.0
    n = int(round(T / dt))
    print(f"markov-synapse network fp{precision}, {n} steps", flush=True)
    net = _markov_net(precision)
    res = net.run(dt=dt, duration=T, spike_recording="population")
    V, G, S = _oracle(net, dt, n, pool_probes=[("I", "ampa"), ("I", "nmda"), ("J", "gaba")])
    for name in net.populations:
        got, ref = res.traces[name]["v"], V[name][:, :, 0]
        err = np.abs(got - ref).max()
        worst = max(rmse(got[:, i], ref[:, i]) for i in range(ref.shape[1]))
        print(name, "max |dV|", err, "worst RMSE", worst, "spikes", int(S[name].sum()))
        if precision == 64:
            assert err < 1e-7
            assert np.array_equal(res.spikes[name], S[name])
        else:
            assert worst <= 0.05
            for i in range(ref.shape[1]):
                a, b = np.flatnonzero(res.spikes[name][:, i]), np.flatnonzero(S[name][:, i])
                assert len(a) == len(b) and (len(a) == 0 or np.max(np.abs(a - b)) <= 1)
    assert S["I"].any()  # AMPA / NMDA input makes the I cells fire, and every pool is driven
    assert min(float(G[k].max()) for k in G) > 0.3


def _projection_net(precision, explicit):
    """Edge sets and projections (``network/edges.py``, ``projections.py``):
    all_pairs, probability and pairs edge sets expanded by per_edge / by_post
    / explicit contacts with per-edge weights and callable delays; with
    ``explicit`` the same contacts are written as plain connections."""
    from braincell_amd.mech import Synapse
    from braincell_amd.network import (Connection, Network, all_pairs, by_post, explicit_contacts, pairs, per_edge,
                                       probability)

    net = Network()
    net.add_population("E1", _cell(3, 0.25, precision=precision))
    net.add_population("E2", _cell(2, 0.6, precision=precision))
    net.add_population("I", _cell(4, 0.0, dend=True, precision=precision,
                                  syn=Synapse("Exp2Syn", tau1=0.5, tau2=4.0, e=0.0, weight=0.01, name="ampa")))
    net.add_population("J", _cell(2, 0.02, precision=precision,
                                  syn=Synapse("ExpSyn", tau=3.0, e=-80.0, weight=0.02, name="gaba")))
    net.add_edges(name="e1i", pre="E1", post="I", method=all_pairs(pre_indices=[0, 2], post_indices=[0, 1, 3]))
    net.add_edges(name="e2i", pre="E2", post="I", method=probability(p=0.6, seed=3))
    net.add_edges(name="ij", pre="I", post="J", method=pairs([(0, 0), (1, 0), (2, 1), (3, 1)]))
    net.project(name="p1", edges="e1i", synapse="ampa", method=per_edge(number=2, replace=False, seed=5),
                weight=np.linspace(0.002, 0.008, 6), delay=lambda c: 0.25 + 0.1 * c.contact_post_index)
    net.project(name="p2", edges="e2i", synapse="ampa", method=by_post(seed=9), weight=0.015, delay=0.3)
    net.project(name="p3", edges="ij", synapse="gaba", method=explicit_contacts([0, 1, 2, 3, 3]), delay=1.0)
    if not explicit:
        return net
    cons = net.all_connections()
    direct = Network()
    for name, pop in net.populations.items():
        direct.add_population(name, pop.cell)
    for c in cons:
        direct.add_connection(Connection(c.pre_population, c.post_population, c.pre_index, c.post_index,
                                         c.synapse, weight=c.weight, delay=c.delay, synapse_index=c.synapse_index))
    return direct


def test_projection_network_equals_explicit_connections_and_oracle(cuda):
    """The projection-built network and its contacts written as connections
    give bitwise identical traces and spikes; FP64 agrees with the oracle to
    1e-7 mV with identical spike steps."""
    dt, T = 0.025, 15.0
    n = int(round(T / dt))
    print(f"projection network fp64, {n} steps", flush=True)
    net = _projection_net(64, False)
    assert sum(c.n_contact for c in net.all_connections()) > 12
    a = net.run(dt=dt, duration=T, spike_recording="population")
    b = _projection_net(64, True).run(dt=dt, duration=T, spike_recording="population")
    V, _, S = _oracle(net, dt, n)
    for name in net.populations:
        assert np.array_equal(a.traces[name]["v"], b.traces[name]["v"])
        assert np.array_equal(a.spikes[name], b.spikes[name])
        err = np.abs(a.traces[name]["v"] - V[name][:, :, 0]).max()
        print(name, "max |dV|", err, "spikes", int(S[name].sum()))
        assert err < 1e-7
        assert np.array_equal(a.spikes[name], S[name])
    assert S["I"].any()


def test_consecutive_runs_continue_the_network(cuda):
    """A second ``Network.run`` continues cells, synapses and clock from the
    first and drops events in flight (``engine.py:294-366, 602-609``); FP64
    against the oracle run the same way; after ``reset_state`` a run repeats
    the first one exactly."""
    from oracle.network import NetworkOracle

    dt = 0.025
    n1, n2 = 200, 400
    print(f"network continuation fp64, {n1} + {n2} steps", flush=True)
    net = _net(64)
    r1 = net.run(dt=dt, duration=n1 * dt, spike_recording="population")
    r2 = net.run(dt=dt, duration=n2 * dt, spike_recording="population")
    assert r2.time[0] == pytest.approx(n1 * dt) and r2.time.shape == (n2,)
    cons = [dict(pre=c.pre_population, post=c.post_population, pre_index=c.pre_index, post_index=c.post_index,
                 synapse=c.synapse, synapse_index=c.synapse_index, weight=c.weight, delay=c.delay)
            for c in net.all_connections()]
    o = NetworkOracle([(n, p.cell.lower(), p.size, p.cell.V_th) for n, p in net.populations.items()], cons, dt)
    cvs = {n: [o.pops[n]["cells"][0]._locate(0, 0.5)[1]] for n in net.populations}
    V1, _, S1 = o.run(n1, probe_cvs=cvs)
    o.reset_delivery()
    V2, _, S2 = o.run(n2, probe_cvs=cvs, step0=n1)
    for name in net.populations:
        for r, V, S in ((r1, V1, S1), (r2, V2, S2)):
            err = np.abs(r.traces[name]["v"] - V[name][:, :, 0]).max()
            assert err < 1e-7, (name, err)
            assert np.array_equal(r.spikes[name], S[name])
    assert S2["I"].any()
    net.reset_state()
    r3 = net.run(dt=dt, duration=n1 * dt, spike_recording="population")
    for name in net.populations:
        assert np.array_equal(r3.traces[name]["v"], r1.traces[name]["v"])
