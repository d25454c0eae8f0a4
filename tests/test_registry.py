"""Integrator registry contract (restates braincell/quad/_registry_test.py:44-345)."""

import warnings

import pytest

from braincell_amd.quad import all_integrators, get_integrator, get_registry, staggered_mi355x_step
from braincell_amd.quad._registry import IntegratorRegistry


def f(target, *args):
    return None


def g(target, *args):
    return None


def test_register_and_lookup_by_alias():
    r = IntegratorRegistry()
    e = r.register("euler", f, aliases=("explicit",), category="explicit", order=1)
    assert r["euler"] is f and r["explicit"] is f
    assert r.resolve("explicit") == "euler"
    assert e.module == f.__module__
    assert "explicit" in r and "nope" not in r and 3 not in r
    assert r.names() == ["euler"] and r.names(include_aliases=True) == ["euler", "explicit"]


def test_collisions_and_types():
    r = IntegratorRegistry()
    r.register("a", f, aliases=("b",))
    with pytest.raises(ValueError):
        r.register("a", g)
    with pytest.raises(ValueError):
        r.register("b", g)
    with pytest.raises(ValueError):
        r.register("c", g, aliases=("b",))
    with pytest.raises(ValueError):
        r.register("d", g, aliases=("a",))
    with pytest.raises(TypeError):
        r.register("", g)
    with pytest.raises(TypeError):
        r.register("e", 3)
    with pytest.raises(TypeError):
        r.register("e", g, aliases=("",))


def test_override_warns_and_drops_aliases():
    r = IntegratorRegistry()
    r.register("a", f, aliases=("b",))
    with pytest.warns(RuntimeWarning):
        r.register("a", g, override=True)
    assert r["a"] is g and "b" not in r


def test_deprecated_warns_once():
    r = IntegratorRegistry()
    r.register("old", f, deprecated=True)
    with pytest.warns(DeprecationWarning):
        r["old"]
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        r["old"]


def test_unregister_and_categories_and_suggest():
    r = IntegratorRegistry()
    r.register("rk4", f, aliases=("runge_kutta4",), category="explicit")
    r.register("rk2", g, category="explicit")
    assert [e.name for e in r.by_category("explicit")] == ["rk2", "rk4"]
    assert r.suggest("rk5") in (["rk4"], ["rk2"])
    r.unregister("rk4")
    assert "runge_kutta4" not in r
    with pytest.raises(KeyError):
        r.unregister("runge_kutta4")
    assert r.get("zzz", 5) == 5


def test_global_registry_has_the_mi355x_solvers():
    assert get_integrator("staggered_mi355x") is staggered_mi355x_step
    assert get_integrator("stagger_mi355x") is staggered_mi355x_step
    assert get_registry().entry("staggered_mi355x").category == "staggered"
    assert "ind_exp_euler_mi355x" in all_integrators
    assert get_integrator(f) is f
    with pytest.raises(ValueError, match="Did you mean 'staggered_mi355x'"):
        get_integrator("staggered_mi355")
    with pytest.raises(TypeError):
        get_integrator(3)


def test_step_rejects_wrong_target():
    with pytest.raises(TypeError):
        staggered_mi355x_step(object())


def test_voltage_integrator_is_registered():
    from braincell_amd.quad import get_integrator
    from braincell_amd.quad._registry import get_registry

    f = get_integrator("dhs_voltage_mi355x")
    assert callable(f)
    assert get_registry().entry("dhs_voltage_mi355x").category == "voltage"
    with pytest.raises(TypeError):
        f(object(), 0.0, 0.025)


def test_point_implicit_integrators_are_registered():
    from braincell_amd.quad import get_integrator
    from braincell_amd.quad._registry import get_registry
    from braincell_amd.single import BACKWARD_EULER, IMPLICIT_EULER, IND_EXP_EULER, SingleCompartment

    for name, method in (("backward_euler_mi355x", BACKWARD_EULER), ("implicit_euler_mi355x", IMPLICIT_EULER)):
        e = get_registry().entry(name)
        assert e.category == "implicit" and e.order == 1
        sc = SingleCompartment(4, channels={"IL": {"g_max": 1e-4}}, solver=name)
        assert sc.method == method
        with pytest.raises(TypeError):
            get_integrator(name)(object())
    assert SingleCompartment(4, channels={"IL": {"g_max": 1e-4}}).method == IND_EXP_EULER
