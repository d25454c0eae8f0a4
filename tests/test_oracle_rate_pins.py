"""The oracle's rate functions and mechanism math against characteristic
pins of the published rate laws, and a mutation check of those pins: every
nonzero numeric literal of ``oracle/channels.py`` changed by 10 % (one at a
time, in a private copy of the module) must fail at least one pin (CPU only)."""

import ast
import os
import types

import numpy as np
import pytest

from oracle import channels as ch
from tests import rate_pins as RP

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "channels.py")

# Literals whose 10 % change leaves every result unchanged to rounding: the
# switch between the Taylor series and expm1(x)/x in the complex-step exprel
# (both agree to 1e-16 relative at |x| ~ 1e-6), and the placeholder divisor
# that keeps the discarded branch of its np.where finite.
EQUIVALENT = {("_exprel", 1e-06), ("_exprel", 1.0)}


@pytest.mark.parametrize("pin", RP.PINS, ids=lambda p: p.__name__)
def test_oracle_meets_pin(pin):
    pin(ch)


def _literals(tree):
    """(function name or '<module>', node) for every nonzero int/float literal
    outside docstrings."""
    out = []
    for top in tree.body:
        scope = top.name if isinstance(top, ast.FunctionDef) else "<module>"
        for node in ast.walk(top):
            if isinstance(node, ast.Constant) and type(node.value) in (int, float) and node.value != 0:
                out.append((scope, node))
    return out


def _mutants():
    src = open(SRC).read()
    tree = ast.parse(src)
    n = len(_literals(tree))
    for i in range(n):
        t = ast.parse(src)
        scope, node = _literals(t)[i]
        orig = node.value
        node.value = float(orig) * 1.1
        yield scope, orig, node.lineno, compile(t, SRC, "exec")


def test_every_literal_is_pinned():
    survivors = []
    count = 0
    for scope, orig, line, code in _mutants():
        count += 1
        mod = types.ModuleType("oracle_channels_mutant")
        mod.__file__ = SRC
        exec(code, mod.__dict__)
        caught = False
        with np.errstate(all="ignore"):
            for pin in RP.PINS:
                try:
                    pin(mod)
                except (AssertionError, FloatingPointError, ZeroDivisionError, ValueError):
                    caught = True
                    break
        if not caught and (scope, orig) not in EQUIVALENT:
            survivors.append(f"{scope} line {line}: {orig}")
    assert count > 100
    assert not survivors, "literals of oracle/channels.py that no pin catches:\n" + "\n".join(survivors)
