"""Edge sets and projections (host logic): the reference's edge-builder and
projection tests restated (``network/edges_test.py:38-207``,
``network/projections_test.py:40-263``), their network-level delivery checks
run through the network oracle with the reference tests' toy solvers, and
the product's contact tables for a projection-built network against the
same network written as explicit connections."""

import numpy as np
import pytest

from braincell_amd import network as N
from braincell_amd.mech import Synapse
from braincell_amd.network import (ContactTable, EdgeSet, Network, Projection, ProjectionEdgeContext, all_pairs,
                                   by_post, dense, explicit_contacts, pairs, per_edge, probability)
from tests.test_oracle_network import DOWN, UP, _toy, soma_cell

EXP = Synapse("ExpSyn", tau=2.0, e=0.0, weight=1.0, name="exp")


def _cons(net):
    return [dict(pre=c.pre_population, post=c.post_population, pre_index=c.pre_index, post_index=c.post_index,
                 synapse=c.synapse, synapse_index=c.synapse_index, weight=c.weight, delay=c.delay)
            for c in net.all_connections()]


def _run_toy(net, n_steps=2):
    pops = [("E", net.populations["E"].cell, -10.0, UP), ("I", net.populations["I"].cell, -65.0, DOWN)]
    G, _ = _toy(pops, _cons(net), n_steps)
    return G[("I", "exp")]


def _two_pops(n_pre=2, n_post=2, dend=False):
    net = Network(name="demo")
    net.add_population("E", soma_cell(n_pre))
    net.add_population("I", soma_cell(n_post, synapse=EXP, dend=dend))
    return net


# ------------------------------------------------------------- edge sets
def test_edge_set_stores_cell_level_edges_and_has_a_compact_repr():
    es = EdgeSet("E_to_I", "E", "I", [0, 1], [2, 3])
    assert es.n_edge == 2 and es.pre_index.dtype == np.int32
    np.testing.assert_array_equal(es.post_index, [2, 3])
    text = repr(es)
    assert "EdgeSet" in text and "n_edge=2" in text and "array(" not in text
    with pytest.raises(ValueError):
        EdgeSet("bad", "E", "I", [0, 1], [2])
    with pytest.raises(ValueError):
        EdgeSet("", "E", "I", [0], [1])
    assert not hasattr(N, "all_to_all") and not hasattr(N, "build")


def test_builders_give_the_reference_tests_edges():
    net = _two_pops(3, 3)
    explicit = net.add_edges(name="explicit", pre="E", post="I", method=pairs([(2, 1), (0, 2)]))
    dense_es = net.add_edges(name="dense", pre="E", post="I",
                             method=dense([[False, True, False], [True, False, True], [False, False, False]]))
    grid = net.add_edges(name="grid", pre="E", post="I", method=all_pairs(pre_indices=[0, 2], post_indices=[1, 2]))
    a = net.add_edges(name="a", pre="E", post="I", method=probability(p=0.5, seed=11))
    b = net.add_edges(name="b", pre="E", post="I", method=probability(p=0.5, seed=11))
    np.testing.assert_array_equal(explicit.pre_index, [2, 0])
    np.testing.assert_array_equal(explicit.post_index, [1, 2])
    np.testing.assert_array_equal(dense_es.pre_index, [0, 1, 1])
    np.testing.assert_array_equal(dense_es.post_index, [1, 0, 2])
    np.testing.assert_array_equal(grid.pre_index, [0, 0, 2, 2])
    np.testing.assert_array_equal(grid.post_index, [1, 2, 1, 2])
    np.testing.assert_array_equal(a.pre_index, b.pre_index)
    np.testing.assert_array_equal(a.post_index, b.post_index)
    assert set(net.edge_sets) == {"explicit", "dense", "grid", "a", "b"}


def test_probability_and_all_pairs_limits():
    net = _two_pops(4, 4)
    full = net.add_edges(name="full", pre="E", post="I", method=probability(p=1.0, seed=0))
    none = net.add_edges(name="none", pre="E", post="I", method=probability(p=0.0, seed=0))
    noself = net.add_edges(name="noself", pre="E", post="I", method=probability(p=1.0, seed=0, allow_self=False))
    ap = net.add_edges(name="ap", pre="E", post="I", method=all_pairs(allow_self=False))
    assert full.n_edge == 16 and none.n_edge == 0
    assert noself.n_edge == 12 and np.all(noself.pre_index != noself.post_index)
    np.testing.assert_array_equal(np.c_[ap.pre_index, ap.post_index], np.c_[noself.pre_index, noself.post_index])
    with pytest.raises(ValueError):
        probability(p=1.5)
    with pytest.raises(ValueError):
        net.add_edges(name="x", pre="E", post="I", method=all_pairs(pre_indices="some"))


def test_custom_callables_are_validated():
    net = _two_pops(3, 3)

    def ring(*, n_pre, n_post):
        idx = list(range(n_pre))
        return idx, [(i + 1) % n_post for i in idx]

    es = net.add_edges(name="ring", pre="E", post="I", method=ring)
    assert es.pre_index.dtype == np.int32 and es.post_index.dtype == np.int32
    np.testing.assert_array_equal(es.post_index, [1, 2, 0])
    with pytest.raises(IndexError, match="pre_index out of range"):
        net.add_edges(name="bad", pre="E", post="I", method=lambda *, n_pre, n_post: ([0, 3], [0, 1]))
    with pytest.raises(ValueError, match="pair"):
        net.add_edges(name="bad2", pre="E", post="I", method=lambda *, n_pre, n_post: [0])
    with pytest.raises(TypeError):
        net.add_edges(name="bad3", pre="E", post="I", method=3)
    with pytest.raises(KeyError):
        net.add_edges(name="bad4", pre="X", post="I", method=ring)
    with pytest.raises(ValueError):
        net.add_edges(name="ring", pre="E", post="I", method=ring)
    with pytest.raises(ValueError):
        net.make_edges(name="m", pre="E", post="I", connector="star")
    with pytest.raises(TypeError):
        net.make_edges(name="m", pre="E", post="I", connector="pairs", pairs=[(0, 0)], extra=1)
    m = net.make_edges(name="m", pre="E", post="I", connector="pairs", pairs=[(0, 0), (1, 2)])
    np.testing.assert_array_equal(m.post_index, [0, 2])


# ------------------------------------------------------------ projections
def _ctx(pre, post, pool):
    e = np.arange(len(pre), dtype=np.int32)
    return ProjectionEdgeContext(edge_index=e, edge_pre_index=np.asarray(pre, dtype=np.int32),
                                 edge_post_index=np.asarray(post, dtype=np.int32), pre_size=max(pre) + 1,
                                 post_size=max(post) + 1, pool_size=pool, synapse="exp")


def test_explicit_contacts_expand_edges():
    es = EdgeSet("E_to_I", "E", "I", [0, 1], [2, 3])
    proj = Projection(name="E_to_I_ampa", edges="E_to_I", synapse="ampa",
                      method=explicit_contacts(source_edge=[0, 0, 1], synapse_index=[0, 1, 1]),
                      weight=np.array([0.1, 0.2, 0.3]))
    (con,) = proj.to_connections(es, pool_size=2)
    assert con.synapse == "ampa"
    np.testing.assert_array_equal(con.pre_index, [0, 0, 1])
    np.testing.assert_array_equal(con.post_index, [2, 2, 3])
    np.testing.assert_array_equal(con.synapse_index, [0, 1, 1])
    np.testing.assert_allclose(con.weight, [0.1, 0.2, 0.3])
    table = explicit_contacts(source_edge=[0, 0, 1], synapse_index=[0, 1, 1])(_ctx([0, 1], [2, 3], 2))
    assert repr(table) == "ContactTable(n_contact=3)"
    text = repr(Projection(name="p", edges="E_to_I", synapse="ampa", edge_index=[0], weight=0.1,
                           method=explicit_contacts([0], [0])))
    assert "edges='E_to_I'" in text and "edge_index=True" in text and "<function" not in text
    with pytest.raises(IndexError):
        Projection(name="p", edges=es, synapse="ampa", method=explicit_contacts([0], [2])).to_connections(
            es, pool_size=2)
    with pytest.raises(IndexError):
        Projection(name="p", edges=es, synapse="ampa", edge_index=[2]).to_connections(es)


def test_per_edge_without_replacement_covers_the_pool_per_edge():
    es = EdgeSet("E_to_I", "E", "I", [0, 1], [1, 1])
    proj = Projection(name="p", edges="E_to_I", synapse="exp", method=per_edge(number=2, replace=False, seed=1),
                      weight=np.array([0.1, 0.2]))
    (con,) = proj.to_connections(es, pool_size=2)
    np.testing.assert_array_equal(con.pre_index, [0, 0, 1, 1])
    np.testing.assert_array_equal(con.post_index, [1, 1, 1, 1])
    np.testing.assert_allclose(con.weight, [0.1, 0.1, 0.2, 0.2])
    for s in (0, 2):
        assert set(con.synapse_index[s:s + 2].tolist()) == {0, 1}
    with pytest.raises(ValueError, match="pool_size"):
        Projection(name="q", edges="E_to_I", synapse="exp",
                   method=per_edge(number=3, replace=False)).to_connections(es, pool_size=2)
    zero = Projection(name="z", edges="E_to_I", synapse="exp", method=per_edge(number=0))
    assert zero.to_connections(es, pool_size=2)[0].n_contact == 0


def test_by_post_draws_jointly_per_postsynaptic_cell():
    es = EdgeSet("E_to_I", "E", "I", [0, 1, 0], [1, 1, 0])
    (con,) = Projection(name="p", edges="E_to_I", synapse="exp",
                        method=by_post(replace=False, seed=1)).to_connections(es, pool_size=2)
    assert set(con.synapse_index[con.post_index == 1].tolist()) == {0, 1}
    with pytest.raises(ValueError, match="total contacts per post"):
        Projection(name="q", edges="E_to_I", synapse="exp",
                   method=by_post(number=2, replace=False, seed=1)).to_connections(es, pool_size=3)


def test_callable_number_weight_and_custom_contact_method():
    es = EdgeSet("E_to_I", "E", "I", [0, 1, 2, 3], [0, 0, 1, 1])
    proj = Projection(name="p", edges="E_to_I", synapse="exp",
                      method=per_edge(number=lambda c: np.where(c.edge_pre_index < c.pre_size // 2, 1, 2),
                                      replace=True, seed=1),
                      weight=lambda c: np.where(c.contact_pre_index < c.pre_size // 2, 0.1, 0.3))
    (con,) = proj.to_connections(es, pre_size=4, post_size=2, pool_size=3)
    np.testing.assert_array_equal(con.pre_index, [0, 1, 2, 2, 3, 3])
    np.testing.assert_allclose(con.weight, [0.1, 0.1, 0.3, 0.3, 0.3, 0.3])

    es2 = EdgeSet("E_to_I", "E", "I", [0, 1], [1, 1])
    seen = []

    def dup_first(ctx):
        seen.append(ctx.synapse)
        return ContactTable(source_edge=[0, 0, 1], synapse_index=[0, 1, 0])

    (con,) = Projection(name="c", edges="E_to_I", synapse="exp", method=dup_first,
                        weight=np.array([0.1, 0.2, 0.3])).to_connections(es2, pool_size=2)
    assert seen == ["exp"]
    np.testing.assert_array_equal(con.pre_index, [0, 0, 1])
    np.testing.assert_array_equal(con.synapse_index, [0, 1, 0])
    np.testing.assert_allclose(con.weight, [0.1, 0.2, 0.3])
    with pytest.raises(TypeError):
        Projection(name="c", edges="E_to_I", synapse="exp", method=lambda ctx: 1).to_connections(es2)
    with pytest.raises(ValueError):
        Projection(name="c", edges="E_to_I", synapse="exp", weight=np.ones(5)).to_connections(es2)
    with pytest.raises(TypeError):
        Projection(name="c", edges=3, synapse="exp")


def test_add_projection_forms_and_errors():
    net = _two_pops()
    net.add_edges(name="E_to_I", pre="E", post="I", method=pairs([(0, 1)]))
    p = net.add_projection(name="a", edges="E_to_I", synapse="exp", weight=0.4)
    assert net.proj["a"] is p and net.projections["a"] is p
    q = net.project(name="b", edges="E_to_I", synapse="exp")
    assert net.proj["b"] is q
    with pytest.raises(ValueError):
        net.add_projection(name="c", edges="E_to_I")
    with pytest.raises(TypeError):
        net.add_projection(Projection(name="d", edges="E_to_I", synapse="exp"), name="d")
    with pytest.raises(ValueError):
        net.add_projection(name="a", edges="E_to_I", synapse="exp")
    net.add_projection(name="ghost", edges="nope", synapse="exp")
    with pytest.raises(KeyError):
        net.all_connections()


# ------------------------------------------- delivery through the oracle
def test_add_edges_and_project_deliver_the_weight():
    """``edges_test.py:91-118``: one edge E0 -> I1, weight 0.5: I1's ExpSyn
    conductance is 0.5 right after the arrival step, I0's stays 0."""
    net = _two_pops()
    net.add_edges(name="E_to_I", pre="E", post="I", method=all_pairs(pre_indices=[0], post_indices=[1]))
    net.project(name="E_to_I_exp", edges="E_to_I", synapse="exp", weight=0.5)
    g = _run_toy(net)
    assert g[0, 1, 0] == 0.0 and g[1, 1, 0] == pytest.approx(0.5, abs=1e-12)
    assert np.all(g[:, 0] == 0.0)


def test_unused_edge_set_delivers_nothing_and_shared_synapse_sums():
    net = _two_pops()
    net.add_edge_set(EdgeSet("unused", "E", "I", [0], [1]))
    assert np.all(_run_toy(net) == 0.0)
    net = _two_pops()
    net.add_edge_set(EdgeSet("E_to_I", "E", "I", [0, 1], [1, 1]))
    net.add_projection(Projection("E_to_I_exp", edges="E_to_I", synapse="exp", weight=np.array([0.25, 0.75])))
    g = _run_toy(net)
    assert g[0, 1, 0] == 0.0
    direct = _two_pops()
    direct.add_connection(N.Connection("E", "I", [0, 1], [1, 1], "exp", weight=[0.25, 0.75]))
    np.testing.assert_array_equal(g, _run_toy(direct))
    assert g[1, 1, 0] == pytest.approx(1.0, abs=1e-12)


def test_projection_onto_pool_locations_matches_explicit_connection_tables():
    """A per_edge projection over a two-location pool lowers to the same
    contact tables (targets, presynaptic cells, delays, weights) as the
    explicit connection holding its contacts."""
    net = _two_pops(3, 2, dend=True)
    net.add_edges(name="E_to_I", pre="E", post="I", method=all_pairs())
    net.add_projection(name="p", edges="E_to_I", synapse="exp", method=per_edge(number=2, replace=False, seed=7),
                       weight=lambda c: 0.1 + 0.01 * c.contact_pre_index, delay=np.array([0.1, 0.2, 0.3, 0.4,
                                                                                           0.5, 0.6]))
    (con,) = net._projection_connections()
    direct = _two_pops(3, 2, dend=True)
    direct.add_connection(N.Connection("E", "I", con.pre_index, con.post_index, "exp", weight=con.weight,
                                       delay=con.delay, synapse_index=con.synapse_index))
    a, b = net.build(0.1), direct.build(0.1)
    for k in ("con_off", "con_pre", "con_delay", "con_w", "syn_row", "syn_row_off", "syn_pack_off"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert a.n_syn == 4 and int(a.con_off[-1]) == 12
    np.testing.assert_array_equal(_run_toy(net, 4), _run_toy(direct, 4))


def test_summaries():
    net = _two_pops(3, 2)
    net.add_edges(name="E_to_I", pre="E", post="I", method=pairs([(0, 1), (2, 0)]))
    net.project(name="p", edges="E_to_I", synapse="exp")
    net.add_connection(N.Connection("E", "I", [1], [1], "exp", weight=0.2))
    assert repr(net) == "Network(name='demo', populations=2, edge_sets=1, projections=1, connections=1)"
    text = str(net)
    for part in ("E: size=3", "E_to_I: E -> I, n_edge=2", "p: edges='E_to_I', synapse='exp'",
                 "[0]: E -> I, synapse='exp', n_contact=1"):
        assert part in text
    assert "<none>" in str(N.Network())
    assert repr(net.populations["I"]).startswith("Population(name='I', size=2")
    con = net.all_connections()[1]
    assert con.n_edge == con.n_contact == 2 and "n_contact=2" in repr(con)
