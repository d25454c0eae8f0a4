"""Name -> step-function registry for integrators.

Same contract as braincell's integrator registry
(``braincell/quad/_registry.py:60-406``): canonical names plus aliases,
collision errors, ``override=True`` with a ``RuntimeWarning``, deprecation
warnings once per name, categories and close-match suggestions.  The
contract is kept identical so that code written against the reference
(``register_integrator`` / ``get_integrator``) selects the MI355X solvers by
name.
"""

from __future__ import annotations

import difflib
import warnings
from dataclasses import dataclass
from typing import Callable, Iterable, Iterator

__all__ = ["IntegratorEntry", "IntegratorRegistry", "get_registry", "register_integrator"]


@dataclass(frozen=True)
class IntegratorEntry:
    name: str
    func: Callable
    aliases: tuple = ()
    category: str = "general"
    order: int | None = None
    description: str = ""
    deprecated: bool = False
    module: str = ""


class IntegratorRegistry:
    """Canonical entries plus an alias index; lookups accept either."""

    def __init__(self) -> None:
        self._entries: dict = {}
        self._aliases: dict = {}
        self._warned: set = set()

    def register(self, name: str, func: Callable, *, aliases: Iterable[str] = (), category: str = "general",
                 order: int | None = None, description: str = "", deprecated: bool = False,
                 override: bool = False) -> IntegratorEntry:
        if not isinstance(name, str) or not name:
            raise TypeError(f"Integrator name must be a non-empty string, got {name!r}.")
        if not callable(func):
            raise TypeError(f"Integrator {name!r} must be a callable, got {type(func).__name__}.")
        aliases = tuple(aliases)
        if any(not isinstance(a, str) or not a for a in aliases):
            raise TypeError(f"Aliases for integrator {name!r} must be non-empty strings, got {aliases!r}.")
        if name in self._entries and not override:
            raise ValueError(f"Integrator {name!r} is already registered by {self._entries[name].module!r}. "
                             "Pass override=True to replace it.")
        owner = self._aliases.get(name)
        if owner is not None and owner != name:
            raise ValueError(f"Cannot register integrator {name!r}: the name is already an alias for {owner!r}.")
        for a in aliases:
            if a in self._entries and a != name:
                raise ValueError(f"Cannot register alias {a!r} for {name!r}: a canonical integrator with that "
                                 "name already exists.")
            o = self._aliases.get(a)
            if o is not None and o != name:
                raise ValueError(f"Cannot register alias {a!r} for {name!r}: it is already an alias for {o!r}.")
        if override and name in self._entries:
            warnings.warn(f"Overriding integrator {name!r} previously registered by "
                          f"{self._entries[name].module!r}.", RuntimeWarning, stacklevel=2)
            self._drop_aliases(name)
        entry = IntegratorEntry(name, func, aliases, category, order, description, deprecated,
                                getattr(func, "__module__", "") or "")
        self._entries[name] = entry
        for a in aliases:
            self._aliases[a] = name
        self._warned.discard(name)
        return entry

    def _drop_aliases(self, name):
        for a in [a for a, t in self._aliases.items() if t == name]:
            del self._aliases[a]

    def unregister(self, name: str) -> None:
        if name not in self._entries:
            raise KeyError(f"No integrator registered as {name!r}.")
        self._drop_aliases(name)
        del self._entries[name]
        self._warned.discard(name)

    def __contains__(self, name: object) -> bool:
        return isinstance(name, str) and (name in self._entries or name in self._aliases)

    def __iter__(self) -> Iterator[str]:
        return iter(self._entries)

    def __len__(self) -> int:
        return len(self._entries)

    def __getitem__(self, name: str) -> Callable:
        return self.entry(name).func

    def get(self, name: str, default=None):
        try:
            return self[name]
        except KeyError:
            return default

    def entry(self, name: str) -> IntegratorEntry:
        canonical = self._aliases.get(name, name)
        try:
            e = self._entries[canonical]
        except KeyError as exc:
            raise KeyError(name) from exc
        if e.deprecated and name not in self._warned:
            self._warned.add(name)
            warnings.warn(f"Integrator {name!r} is deprecated.", DeprecationWarning, stacklevel=3)
        return e

    def resolve(self, name: str) -> str:
        return self.entry(name).name

    def names(self, *, include_aliases: bool = False) -> list:
        out = set(self._entries)
        if include_aliases:
            out |= set(self._aliases)
        return sorted(out)

    def by_category(self, category: str) -> list:
        return sorted((e for e in self._entries.values() if e.category == category), key=lambda e: e.name)

    def items(self):
        for n, e in self._entries.items():
            yield n, e.func

    def entries(self):
        return iter(self._entries.values())

    def as_dict(self, *, include_aliases: bool = True) -> dict:
        out = {n: e.func for n, e in self._entries.items()}
        if include_aliases:
            out.update({a: self._entries[c].func for a, c in self._aliases.items()})
        return out

    def suggest(self, name: str, *, n: int = 1, cutoff: float = 0.6) -> list:
        return difflib.get_close_matches(name, self.names(include_aliases=True), n=n, cutoff=cutoff)


_REGISTRY = IntegratorRegistry()


def get_registry() -> IntegratorRegistry:
    return _REGISTRY


def register_integrator(name: str, *, aliases: Iterable[str] = (), category: str = "general",
                        order: int | None = None, description: str = "", deprecated: bool = False,
                        override: bool = False):
    """Decorator registering a step function in the global registry; the
    function is returned unchanged."""

    def wrap(func):
        _REGISTRY.register(name, func, aliases=aliases, category=category, order=order, description=description,
                           deprecated=deprecated, override=override)
        return func

    return wrap
