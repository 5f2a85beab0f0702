"""The reference's 3- and 4-stage integrator coefficients and stage order
(tests/golden/integrator_coefficients.json, read from hmclab/Samplers.py by
tests/golden/make_integrator_coefficients.py) against the oracle's tables
and the engine's device schedule (hmc_integrator_schedule, host only), and a
planted mistake (a1 and b1 swapped, or a digit slip) that the comparison
must catch: palindromic, consistent schedules pass the integrator invariants
(reversibility, second order), so only the literal values pin them."""
import ast
import json
import os

import numpy as np
import pytest

from oracle import hmc_oracle as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "integrator_coefficients.json")


def _eval(expr, values):
    """Arithmetic of numbers and names only."""
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant):
            return float(node.value)
        if isinstance(node, ast.Name):
            return values[node.id]
        if isinstance(node, ast.BinOp):
            a, b = ev(node.left), ev(node.right)
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b}[type(node.op)]
        raise ValueError(expr)
    return ev(ast.parse(expr, mode="eval"))


def _reference_step(kind, mutate=None):
    fx = json.load(open(FIXTURE))[kind]
    vals = {k: float(v) for k, v in fx["literal"].items()}
    if mutate:
        mutate(vals)
    for k, e in fx["derived"].items():
        vals[k] = _eval(e, vals)
    assert fx["loop_order"] == [s.strip() for s in fx["schema"]]
    return [(name[0], vals[name]) for name in fx["loop_order"]]


def _oracle_step(kind):
    return [(t, c) for t, c in O._stages(kind)]


def _device_events(kind, L):
    from hmclab_amd.engine import integrator_schedule

    first, st = integrator_schedule(kind, L)
    ev = [("a", first)]
    for kick, drift, drift2 in st:
        ev += [("b", kick), ("a", drift)]
        if drift2 != 0.0:
            ev.append(("a", drift2))
    return ev


@pytest.mark.parametrize("kind", ["3s", "4s"])
def test_oracle_tables_equal_the_reference_literals(kind):
    assert _oracle_step(kind) == _reference_step(kind)


@pytest.mark.parametrize("kind", ["3s", "4s"])
@pytest.mark.parametrize("L", [1, 2, 3])
def test_device_schedule_equals_the_reference_literals(kind, L):
    assert _device_events(kind, L) == _reference_step(kind) * L


@pytest.mark.parametrize("L", [1, 2, 5])
def test_device_leapfrog_schedule(L):
    # Samplers.py:1539-1581: half drift, (L - 1) x (kick, drift), kick, half drift
    want = [("a", 0.5)] + [("b", 1.0), ("a", 1.0)] * (L - 1) + [("b", 1.0), ("a", 0.5)]
    assert _device_events("lf", L) == want


def _swap_a1_b1(v):
    v["a1"], v["b1"] = v["b1"], v["a1"]


def _digit_slip(v):
    v["b1"] = float(str(v["b1"]).replace("1916678", "1916687"))


@pytest.mark.parametrize("kind,mutate", [("3s", _swap_a1_b1), ("4s", _swap_a1_b1), ("4s", _digit_slip)])
def test_planted_mistakes_are_caught(kind, mutate):
    wrong = _reference_step(kind, mutate)
    assert wrong != _reference_step(kind)
    assert _oracle_step(kind) != wrong
    assert _device_events(kind, 2) != wrong * 2
