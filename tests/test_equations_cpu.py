"""Equation front end on the host: multi-line if-then-else statements
(parser/ITE.py, StringManipulation.py:153-245) and the logical operators of
conditions (Equation.py:16-27)."""

import pytest

from annarchy_amd import codegen
from annarchy_amd import equations as E


def test_continuation_lines_join_an_if_then_else():
    eqs = E._equations("""
        r = if (t <= 5):
                0
            else:
                if (t >= 7): 2 else: 1
        v = v + 1 : init = 2.0, min = 0
    """)
    assert eqs[0] == ("r = if (t <= 5): 0 else: if (t >= 7): 2 else: 1", "")
    assert eqs[1] == ("v = v + 1", "init = 2.0, min = 0")


def test_ite_tree():
    eq, tree = E._extract_ite("r = if (t <= 5): 0 else: if (t >= 7): 2 else: 1")
    assert eq == "r = __conditional__0"
    assert tree == ["(t <= 5)", "0", ["(t >= 7)", "2", "1"]]
    assert E._extract_ite("r = r + 1") == ("r = r + 1", None)


@pytest.mark.parametrize("bad", ["r = if t > 1: 2", "r = 1 + if t > 1: 2 else: 3"])
def test_ite_errors(bad):
    with pytest.raises(E.EquationError):
        E._extract_ite(bad)


def _line(equations, needle):
    src = codegen.generate_from_text("n", "", equations)
    return [ln.strip() for ln in src.source.split("\n") if needle in ln][0]


def test_ite_code():
    # the kernels name a model's attributes u_<name>
    assert _line("r = if (t is 1): 1.0 else: 0.0", "u_r[i] =") == "u_r[i] = ((double(t)*dt) == 1 ? 1.0 : 0.0);"
    assert _line("dr/dt = if t > 5: 1 else: 0", "double _r =") == "double _r = ((double(t)*dt) > 5 ? 1 : 0);"


def test_logical_operators():
    code = _line("dr/dt = ite((t > 5) and (t < 20), 1, 0)", "double _r =")
    assert "&&" in code and "< 20" in code
    code = _line("r = ite((t > 5) or not (t > 8), 1, 0)", "u_r[i] =")
    assert "||" in code and "<= 8" in code
