"""Step context (``t``, ``dt``, ``precision``) read by the step functions.

Counterpart of the ``brainstate.environ`` values the reference step reads
(``braincell/quad/_staggered.py:130-131``): a thread-local stack of dicts.
"""

from __future__ import annotations

import contextlib
import threading

__all__ = ["context", "get", "get_dt", "set"]

_local = threading.local()
_DEFAULTS = {"precision": 32}


def _stack():
    if not hasattr(_local, "stack"):
        _local.stack = [dict(_DEFAULTS)]
    return _local.stack


@contextlib.contextmanager
def context(**kw):
    st = _stack()
    st.append({**st[-1], **kw})
    try:
        yield
    finally:
        st.pop()


def set(**kw):  # noqa: A001 - mirrors brainstate.environ.set
    _stack()[-1].update(kw)


def get(key, default=None):
    v = _stack()[-1].get(key, default)
    return v


def get_dt():
    dt = get("dt")
    if dt is None:
        raise KeyError("dt is not set: use braincell_amd.environ.context(dt=...)")
    return float(dt)
