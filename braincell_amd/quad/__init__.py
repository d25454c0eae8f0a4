"""Integrators registered for the MI355X kernels.

``get_integrator`` / ``register_integrator`` follow
``braincell/quad/__init__.py:167-214`` and ``quad/_registry.py:355-406``.
Registered here:

- ``"staggered_mi355x"`` (alias ``"stagger_mi355x"``): the staggered step
  (``braincell/quad/_staggered.py:47-154``) on the persistent HIP kernel;
- ``"ind_exp_euler_mi355x"``: independent exp-Euler for point neurons
  (``braincell/quad/_exp_euler.py:217-320``) on the point-neuron kernel;
- ``"backward_euler_mi355x"`` / ``"implicit_euler_mi355x"``: the linearly
  implicit point-neuron steps (``quad/_backward_euler.py:31-160``,
  ``quad/_implicit.py:43-120, 313-368``) on the same kernel file;
- ``"dhs_voltage_mi355x"``: the implicit-Euler voltage solve alone
  (``quad/_staggered.py:187-320``).
"""

from __future__ import annotations

from typing import Callable, Mapping

from .. import environ
from ._registry import IntegratorEntry, IntegratorRegistry, get_registry, register_integrator

__all__ = [
    "IntegratorEntry", "IntegratorRegistry", "get_registry", "register_integrator", "get_integrator",
    "all_integrators", "staggered_mi355x_step", "ind_exp_euler_mi355x_step", "dhs_voltage_mi355x_step",
    "backward_euler_mi355x_step", "implicit_euler_mi355x_step",
]


class _RegistryView(Mapping):
    def __init__(self, reg):
        self._r = reg

    def __getitem__(self, k):
        return self._r[k]

    def __iter__(self):
        return iter(self._r.as_dict(include_aliases=True))

    def __len__(self):
        return len(self._r.as_dict(include_aliases=True))

    def __contains__(self, k):
        return k in self._r


all_integrators: Mapping = _RegistryView(get_registry())


def get_integrator(method) -> Callable:
    """Resolve a step function from a registered name (or alias) or pass a
    callable through; unknown names raise ``ValueError`` with a suggestion,
    other types ``TypeError``."""
    if callable(method):
        return method
    if isinstance(method, str):
        reg = get_registry()
        try:
            return reg[method]
        except KeyError:
            sug = reg.suggest(method, n=1)
            hint = f" Did you mean {sug[0]!r}?" if sug else ""
            raise ValueError(f"Unknown integrator {method!r}.{hint} Available: {', '.join(reg.names())}.")
    raise TypeError(f"Integrator method must be a string or callable, got {type(method).__name__}.")


@register_integrator("staggered_mi355x", aliases=("stagger_mi355x",), category="staggered",
                     description="Staggered DHS voltage / ind_exp_euler channel step on MI355X HIP kernels.")
def staggered_mi355x_step(target, *args):
    """Advance a multi-compartment cell by one staggered step (``t``/``dt``
    from :mod:`braincell_amd.environ`); mutates the cell's device state."""
    from ..cell import Cell

    if not isinstance(target, Cell):
        raise TypeError(f"The stagger integrator only supports Cell, but we got {type(target)} instead.")
    order = target.ion_channel_update_order
    if order not in ("family", "integration"):
        raise ValueError(f"ion_channel_update_order must be 'family' or 'integration', got {order!r}.")
    t = environ.get("t", 0.0)
    dt = environ.get_dt()
    return target.advance(t, dt, 1)


@register_integrator("dhs_voltage_mi355x", category="voltage",
                     description="Implicit-Euler DHS voltage step alone on the MI355X staggered kernel.")
def dhs_voltage_mi355x_step(target, t, dt, *args):
    """``dhs_voltage`` (``quad/_staggered.py:187-320``): one implicit-Euler
    voltage solve of a multi-compartment cell at time ``t`` (ms) with step
    ``dt`` (ms); gates and the Ca pool are left as they are."""
    from ..cell import Cell

    if not isinstance(target, Cell):
        raise TypeError(f"dhs_voltage_step(...) requires a node-tree aware target, got {type(target)}.")
    return target.advance(float(t), float(dt), 1, voltage_only=True)


@register_integrator("ind_exp_euler_mi355x", category="exponential",
                     description="Independent exponential Euler for point neurons on the MI355X HIP kernel.")
def ind_exp_euler_mi355x_step(target, *args):
    """Advance a :class:`~braincell_amd.single.SingleCompartment` population
    by one step with the (Jacobi) independent exp-Euler rule."""
    from ..single import SingleCompartment

    if not isinstance(target, SingleCompartment):
        raise TypeError(f"ind_exp_euler_mi355x supports SingleCompartment, got {type(target)}.")
    t = environ.get("t", 0.0)
    dt = environ.get_dt()
    return target.advance(t, dt, 1, *args)


ind_exp_euler_mi355x_step.point_method = 0


def _point_step(target, name, method, args):
    from ..single import SingleCompartment

    if not isinstance(target, SingleCompartment):
        raise TypeError(f"{name} supports SingleCompartment, got {type(target)}.")
    t = environ.get("t", 0.0)
    dt = environ.get_dt()
    return target.advance(t, dt, 1, *args, method=method)


@register_integrator("backward_euler_mi355x", category="implicit", order=1,
                     description="Linearised backward Euler for point neurons on the MI355X HIP kernel.")
def backward_euler_mi355x_step(target, *args):
    """One step of ``backward_euler`` (``quad/_backward_euler.py:31-160``):
    ``(I - dt J) dy = dt f(y_n)``.  A
    :class:`~braincell_amd.single.SingleCompartment` neuron uses the full
    Jacobian of its 11 states; a passive :class:`~braincell_amd.cell.Cell`
    steps V alone with the column-sum Jacobian of the stacked-state form
    (``braincell_amd.cable``); an active Cell raises ``ValueError`` as the
    reference's state stacking does."""
    from ..cell import Cell

    if isinstance(target, Cell):
        from ..cable import check_passive

        if not target._initialized:
            target.init_state()
        check_passive(target._model)
        return target.advance(environ.get("t", 0.0), environ.get_dt(), 1)
    return _point_step(target, "backward_euler_mi355x", 1, args)


backward_euler_mi355x_step.point_method = 1


@register_integrator("implicit_euler_mi355x", category="implicit", order=1,
                     description="Newton step of the order-2 implicit residual for point neurons on the MI355X HIP kernel.")
def implicit_euler_mi355x_step(target, *args):
    """One step of ``implicit_euler`` (``quad/_implicit.py:43-120, 313-368``):
    the Newton iterate of y - y_n - dt/2 (f(y_n) + f(y)) = 0 from y_n, i.e.
    ``(I - dt/2 J) dy = dt f(y_n)``.  The reference iterates further only
    while the population-wide residual before an update is below 1e-5; those
    iterates move y by the square of that residual and are not taken.

    A :class:`~braincell_amd.cell.Cell` calls its solver as ``solver(cell)``
    (``cell.py:2374-2377``) while the reference's ``implicit_euler_step``
    needs ``(target, t, dt)``, so a Cell cannot use it: TypeError, as there."""
    from ..cell import Cell

    if isinstance(target, Cell):
        raise TypeError("implicit_euler_step() missing 2 required positional arguments: 't' and 'dt' "
                        "(a Cell calls its solver with the cell alone)")
    return _point_step(target, "implicit_euler_mi355x", 2, args)


implicit_euler_mi355x_step.point_method = 2
