"""Neuron description -> C statements (the equation front end).

Restates the parts of ANNarchy's parser that produce the per-variable C
snippets spliced into the update template:

* parameters / equations / flags: ``parser/Extraction.py:251-560``
  (``: population`` makes an attribute population-wide; ``init=``, ``min=``,
  ``max=``, method flags);
* random distributions ``Uniform(a, b)`` / ``Normal(mu, sigma)`` become the
  local variables ``rand_k`` drawn once per step (``Extraction.py:32-122``);
* ``sum(target)`` becomes ``_sum_<target>[i]`` (``AnalyseNeuron.py:244-250``);
* ODEs: explicit Euler ``_x = <sympy-simplified rhs>; x += dt*_x ;``
  (``parser/Equation.py:276-307``) and exponential Euler with the global step
  size hoisted into a pre-loop value (``Equation.py:559-604``, standardisation
  ``:651-741``); assignments, increments and conditions
  (``Equation.py:742-900``); C text from ``sympy.ccode(precision=8)``
  (``Equation.py:153-197``), with the ``f`` suffix and ``expf``/``powf`` names
  in single precision.

The symbols use the reference's placeholders (``v%(local_index)s``,
``tau%(global_index)s``, ``dt``, ``(double(t)*dt)``) so the emitted text is
the text the reference splices into its loops.
"""

from __future__ import annotations

import re
from dataclasses import dataclass, field

import sympy as sp
from sympy.parsing.sympy_parser import convert_xor, parse_expr, standard_transformations

DISTRIBUTIONS = {"Uniform": 0, "Normal": 1}

_FUNCTIONS = {
    "pos": sp.Function("positive"),
    "positive": sp.Function("positive"),
    "neg": sp.Function("negative"),
    "negative": sp.Function("negative"),
    "clip": sp.Function("clip"),
    "ite": sp.Function("ite", nargs=3),
    "modulo": sp.Function("modulo"),
    "power": sp.Function("power", nargs=2),
    "fabs": sp.Function("fabs"),
    "round": sp.Function("round"),
    "True": sp.Symbol("true"),
    "False": sp.Symbol("false"),
}
_USER_FUNCTIONS = {
    "pos": "positive",
    "positive": "positive",
    "neg": "negative",
    "negative": "negative",
    "modulo": "modulo",
    "fabs": "fabs",
    "clip": "clip",
    "ite": "ite",
    "power": "power",
}


class EquationError(ValueError):
    pass


@dataclass
class Variable:
    name: str
    eq: str = ""
    locality: str = "local"          # local | global
    flags: list = field(default_factory=list)
    bounds: dict = field(default_factory=dict)
    init: float = 0.0
    is_parameter: bool = False
    ctype: str = "double"
    # filled by the analysis
    method: str = "explicit"
    cpp: str = ""                    # ODE increment statement or assignment
    switch: str | None = None        # ODE: x += dt*_x ;
    pre_loop: dict = field(default_factory=dict)
    bound_code: str = ""


@dataclass
class RandomVar:
    name: str
    dist: str
    args: tuple
    locality: str


@dataclass
class NeuronDescription:
    name: str
    spiking: bool
    parameters: list
    variables: list
    randoms: list
    targets: list
    spike_cond: str = ""
    spike_reset: list = field(default_factory=list)
    refractory: float | str | None = None

    @property
    def attributes(self):
        return [p.name for p in self.parameters] + [v.name for v in self.variables]

    def local_names(self):
        return [a.name for a in self.parameters + self.variables if a.locality == "local"]

    def global_names(self):
        return [a.name for a in self.parameters + self.variables if a.locality == "global"]


# ---------------------------------------------------------------------------
# String handling (parser/StringManipulation.py)
# ---------------------------------------------------------------------------

def _lines(text: str | None) -> list[str]:
    if not text:
        return []
    out = []
    for raw in text.replace(";", "\n").split("\n"):
        line = raw.split("#", 1)[0].strip()
        if line:
            out.append(line)
    return out


_CONSTRAINT_KEYWORDS = ("init", "min", "max", "population", "postsynaptic", "projection", "explicit",
                        "implicit", "semiimplicit", "exponential", "midpoint", "rk4", "runge-kutta4", "exact",
                        "event-driven", "unless_refractory", "int", "bool", "float")


def _is_constraint(text: str) -> bool:
    """The text after the last ':' of a line is a constraint when it opens
    with a reserved keyword (StringManipulation.py:7-35, :162-168)."""
    e = "," + text.replace(" ", "") + ","
    return any(re.match(r"([,]+)" + re.escape(k) + r"([=,]+)", e) for k in _CONSTRAINT_KEYWORDS)


def _equations(text: str | None) -> list[tuple[str, str]]:
    """(equation, constraint) of every variable (StringManipulation.py:
    153-245): the ':' of an if-then-else is not a constraint separator, and a
    line without an assignment continues the previous equation (multi-line
    if-then-else statements)."""
    out = []
    for line in _lines(text):
        eq, cons = line, ""
        if ":" in line:
            head, tail = line.rsplit(":", 1)
            if _is_constraint(tail):
                eq, cons = head.strip(), tail.strip()
        if not re.findall(r"([\s\w\+\-\*\/\)]+)=([^=])", eq) and out:
            prev_eq, prev_cons = out[-1]
            out[-1] = (prev_eq + " " + eq.strip(), (prev_cons + ", " + cons) if prev_cons and cons else prev_cons + cons)
        else:
            out.append((eq.strip(), cons))
    return out


_ITE_PH = "__conditional__0"


def _logical(expr: str) -> str:
    """Natural logical operators in sympy form (Equation.py:16-27, applied to
    every expression before its spaces go)."""
    return (expr.replace(" and ", " & ").replace(" or ", " | ").replace(" is not ", " != ")
            .replace(" not ", " Not ").replace(" not(", " Not(").replace(" is ", " == "))


def _extract_ite(eq: str):
    """One if-then-else right of the '=' (ITE.py:88-180): returns the equation
    with the placeholder __conditional__0 and the nested [cond, then, else]
    tree, or (eq, None) without a conditional."""
    if "if " not in eq:
        return eq, None
    left, right = eq.split("=", 1)
    if right.count(":") == 0:
        raise EquationError(f"conditional statements define then and else values: {eq}")
    if not right.strip().startswith("if"):
        raise EquationError(f"the right term must start with the if statement: {eq}")
    if right.count(":") != 2 * len(re.findall("else", right)):
        raise EquationError(f"conditional statements must use both : and else: {eq}")
    tokens = []
    for item in (a.strip() for a in right.split(":")):
        if item.startswith("if "):
            tokens.append(item)
        elif item.endswith("else"):
            tokens.append(item[: -len("else")].strip())
            tokens.append("else")
        else:
            tokens.append(item)

    def tree(toks):
        out = []
        while toks:
            if toks[0].startswith("if"):
                node = [toks.pop(0)[2:].strip(), tree(toks)]
                if toks and toks[0] == "else":
                    toks.pop(0)
                    node.append(tree(toks))
                else:
                    raise EquationError(f"an if without else: {eq}")
                out.append(node)
            elif toks[0] != "else":
                out.append(toks.pop(0))
            else:
                break
        return out[0]

    return left + "= " + _ITE_PH, tree(tokens)


def _ite_code(tr, node) -> str:
    if isinstance(node, list):
        return ("(" + tr.condition(node[0]) + " ? " + _ite_code(tr, node[1]) + " : "
                + _ite_code(tr, node[2]) + ")")
    return tr.expression(node)


def _split_constraint(line: str) -> tuple[str, str]:
    # "eq : flags" -- the last colon that is not inside an ite(...) / brackets
    depth = 0
    for k in range(len(line) - 1, -1, -1):
        c = line[k]
        if c == ")":
            depth += 1
        elif c == "(":
            depth -= 1
        elif c == ":" and depth == 0:
            return line[:k].strip(), line[k + 1:].strip()
    return line.strip(), ""


def _flags(constraint: str) -> tuple[dict, list]:
    bounds, flags = {}, []
    for part in [p.strip() for p in constraint.split(",") if p.strip()]:
        if "=" in part:
            k, v = part.split("=", 1)
            bounds[k.strip()] = v.strip()
        else:
            flags.append(part)
    return bounds, flags


def _variable_name(eq: str) -> str:
    e = eq.replace(" ", "")
    m = re.search(r"d([A-Za-z_]\w*)/dt", e)
    if m:
        return m.group(1)
    for op in ["+=", "-=", "*=", "/="]:
        if op in e:
            return e.split(op, 1)[0]
    lhs = e.split("=", 1)[0]
    return lhs


def _parse_value(v: str) -> float:
    v = v.strip()
    if v in ("True", "true"):
        return 1.0
    if v in ("False", "false"):
        return 0.0
    return float(eval(v, {"__builtins__": {}}, {}))


# ---------------------------------------------------------------------------
# Symbolic translation
# ---------------------------------------------------------------------------

class _Translator:
    def __init__(self, desc: NeuronDescription, precision: str = "double"):
        self.desc = desc
        self.precision = precision
        d = {"dt": sp.Symbol("dt"), "t": sp.Symbol("(double(t)*dt)")}
        d.update(_FUNCTIONS)
        for n in desc.local_names():
            d[n] = sp.Symbol(n + "%(local_index)s")
        for n in desc.global_names():
            d[n] = sp.Symbol(n + "%(global_index)s")
        for r in desc.randoms:
            idx = "%(local_index)s" if r.locality == "local" else "%(global_index)s"
            d[r.name] = sp.Symbol(r.name + idx)
        for tgt in desc.targets:
            d["__sum_" + tgt + "__"] = sp.Symbol("_sum_" + tgt + "%(local_index)s")
        self.local_dict = d

    def parse(self, expr: str, extra: dict | None = None):
        d = dict(self.local_dict)
        d.setdefault("Not", sp.Not)
        if extra:
            d.update(extra)
        expr = _logical(expr)
        try:
            return parse_expr(expr, local_dict=d,
                              transformations=(standard_transformations + (convert_xor,)),
                              evaluate=False)
        except Exception as e:  # pragma: no cover - message path
            raise EquationError(f"cannot analyse the expression: {expr} ({e})") from e

    def ccode(self, e) -> str:
        try:
            code = sp.ccode(e, precision=8, strict=False, user_functions=_USER_FUNCTIONS)
        except TypeError:  # sympy < 1.13
            code = sp.ccode(e, precision=8, user_functions=_USER_FUNCTIONS)
        if self.precision == "float":
            # Equation.py:171-195: f suffix on literals, single-precision functions
            matches = sorted(set(re.findall(r"[-]?[0-9]+\.[0-9]+", code)))
            for m in matches:
                code = code.replace(m, str(round(float(m), 8)) + "f")
            for m in re.findall(r"[-]?[0-9]+\.[0-9]+f[f]+", code):
                code = code.replace(m, str(round(float(m.split("f")[0]), 8)) + "f")
            for fn in ["fabs", "pow", "exp", "sin", "cos", "tan", "round"]:
                code = code.replace(fn + "(", fn + "f(")
        return code

    # -- ODEs ---------------------------------------------------------------
    def explicit(self, name: str, expression: str) -> tuple[dict, str, str]:
        expression = expression.replace("d" + name + "/dt", "_grad_var_")
        g = sp.Symbol("_grad_var_")
        a = self.parse(expression, {"_grad_var_": g})
        eq = sp.solve(a, g, check=False, rational=False)[0]
        eq = sp.simplify(eq, ratio=1.0)
        code = f"{self.precision} _{name} = {self.ccode(eq)};"
        switch = self.ccode(self.local_dict[name]) + " += dt*_" + name + " ;"
        return {}, code, switch

    def standardize(self, name: str, expression: str):
        expression = expression.replace("d" + name + "/dt", "_gradvar_")
        g = sp.Symbol("_gradvar_")
        a = self.parse(expression, {"_gradvar_": g})
        var = self.local_dict[name]
        expanded = a.expand(modulus=None, power_base=False, power_exp=False, mul=True,
                            log=False, multinomial=False)
        coll = sp.collect(expanded, var, evaluate=False, exact=False)
        if var not in coll or len(coll) > 2:
            coll = sp.collect(sp.simplify(expanded), var, evaluate=False, exact=False)
            if var not in coll or len(coll) > 2:
                return None, None, None
        factor_var = coll[var]
        cg = sp.collect(sp.expand(a, g), g, evaluate=False, exact=False)
        factor_grad = cg[g] if g in cg else sp.S(1.0)
        real_tau = factor_grad / factor_var
        normalized = a / factor_var
        steady = sp.simplify(sp.together(real_tau * g + var - normalized))
        stepsize = sp.simplify(sp.together(self.local_dict["dt"] / real_tau))
        return real_tau, stepsize, steady

    def exponential(self, name: str, expression: str, local_names) -> tuple[dict, str, str]:
        real_tau, stepsize, steady = self.standardize(name, expression)
        if real_tau is None:
            return self.explicit(name, expression)
        global_step = not any(self.local_dict[v] in stepsize.atoms() for v in local_names)
        pre = {}
        step_code = "(1.0 - exp(" + self.ccode(-stepsize) + "))"
        if global_step:
            pre = {"name": "__stepsize_" + name, "value": "1.0 - exp( " + self.ccode(-stepsize) + ")"}
            step_code = "__stepsize_" + name
        var = self.ccode(self.local_dict[name])
        code = (f"{self.precision} _{name} =  {step_code}*({self.ccode(steady)} - "
                f"{self.ccode(self.local_dict[name])});")
        switch = var + " += _" + name + " ;"
        return pre, code, switch

    # -- assignments, increments, conditions --------------------------------
    def assignment(self, expression: str) -> str:
        name = expression[: expression.find("=")].strip()
        rhs = expression[expression.find("=") + 1:]
        a = self.parse(rhs)
        return self.ccode(self.local_dict[name]) + " = " + self.ccode(a) + ";"

    def increment(self, expression: str) -> str:
        for op in ["+=", "-=", "*=", "/="]:
            if op in expression:
                name = expression[: expression.find(op)].strip()
                rhs = expression[expression.find(op) + 2:]
                a = self.parse(rhs)
                return (self.ccode(self.local_dict[name]) + f" {op} " +
                        self.ccode(sp.simplify(a, ratio=1.0)) + ";")
        raise EquationError(expression)

    def condition(self, expression: str) -> str:
        expression = expression.strip()
        expression = (expression.replace(" and ", " & ").replace(" or ", " | ")
                      .replace(" is not ", " != ").replace(" not ", " Not ").replace(" is ", " == "))
        if "==" in expression:
            if "&" in expression or "|" in expression:
                expression = re.sub(r"([\w\s.]+)==([\w\s.]+)", r"Equality(\1, \2)", expression)
            else:
                if expression[0] == "(" and expression[-1] == ")":
                    expression = expression[1:-1]
                terms = expression.split("==")
                expression = "Equality(" + terms[0] + ", " + terms[1] + ")"
        if "!=" in expression:
            if "&" in expression or "|" in expression:
                expression = re.sub(r"([\w\s.]+)!=([\w\s.]+)", r"Not(Equality(\1, \2))", expression)
            else:
                if expression[0] == "(" and expression[-1] == ")":
                    expression = expression[1:-1]
                terms = expression.split("!=")
                expression = "Not(Equality(" + terms[0] + ", " + terms[1] + "))"
        extra = {"Equality": sp.Eq, "Not": sp.Not}
        return self.ccode(self.parse(expression, extra))

    def expression(self, expr: str) -> str:
        return self.ccode(self.parse(expr))


def _equation_type(name: str, expression: str) -> str:
    e = expression.replace(" ", "")
    for op in ["+=", "-=", "*=", "/="]:
        if op in e:
            return "inc"
    if "d" + name + "/dt" in e:
        return "ODE"
    split = e.split("=")
    if len(split) == 1 or split[1] == "":
        return "cond"
    return "simple"


# ---------------------------------------------------------------------------
# Public entry point
# ---------------------------------------------------------------------------

def _constant_params(constants, declared, texts, precision: str = "double") -> list:
    """Network constants read by a model (core/Constant.py:16-60): each one
    becomes a model-wide parameter flagged "constant" (initial value 0; the
    network writes the constant's value into it), unless the model declares
    an attribute of the same name, which hides the constant."""
    if not constants:
        return []
    blob = "\n".join(t for t in texts if t)
    out = []
    for cname in sorted(constants):
        if cname in declared:
            continue
        if re.search(r"(?<![\w.])" + re.escape(cname) + r"(?!\w)", blob):
            out.append(Variable(name=cname, eq=f"{cname} = 0.0", locality="global", flags=["constant"],
                                bounds={}, init=0.0, is_parameter=True, ctype=precision))
    return out


def analyse_neuron(parameters: str = "", equations: str = "", spike: str | None = None,
                   reset: str | None = None, refractory=None, name: str = "neuron",
                   precision: str = "double", constants=None, incoming_targets=None) -> NeuronDescription:
    """Analyse a neuron definition into C snippets (AnalyseNeuron.py:27-380).
    `constants`: names of the network constants an equation may read;
    `incoming_targets`: targets of the rate-coded projections into the population,
    each of which gets its sum even when no equation reads it."""
    params = []
    for line in _lines(parameters):
        eq, constraint = _split_constraint(line)
        bounds, flags = _flags(constraint)
        if "=" not in eq:
            raise EquationError(f"parameter without value: {line}")
        pname, val = eq.split("=", 1)
        params.append(Variable(name=pname.strip(), eq=eq, locality="global" if "population" in flags else "local",
                               flags=flags, bounds=bounds, init=_parse_value(val), is_parameter=True,
                               ctype="int" if "int" in flags else precision))
    variables = []
    for eq, constraint in _equations(equations):
        bounds, flags = _flags(constraint)
        vname = _variable_name(eq)
        init = 0.0
        if "init" in bounds:
            init = _parse_value(bounds.pop("init"))
        variables.append(Variable(name=vname, eq=eq, locality="global" if "population" in flags else "local",
                                  flags=flags, bounds=bounds, init=init,
                                  ctype="int" if "int" in flags else precision))
    declared = {v.name for v in params + variables}
    for c in _constant_params(constants, declared | {"r"} if spike is not None else declared,
                              [equations, spike, reset], precision):
        params.append(c)
        declared.add(c.name)
    spiking = spike is not None
    if spiking:
        # spiking neurons hold their mean firing rate in r (0 unless
        # compute_firing_rate() is enabled), AnalyseNeuron.py:145-160
        if "r" in declared:
            raise EquationError('spiking neurons use the variable "r" for the mean firing rate: use another name')
        variables.append(Variable(name="r", eq="", locality="local", flags=[], bounds={}, init=0.0,
                                  ctype=precision))
        declared.add("r")
    # conductances of spiking neurons and targets of rate neurons
    targets = []
    for v in variables:
        for m in re.finditer(r"sum\(\s*(\w+)\s*\)", v.eq):
            if m.group(1) not in targets:
                targets.append(m.group(1))
    if not spiking:
        for tgt in incoming_targets or []:
            if tgt not in targets:
                targets.append(tgt)
    if spiking:
        for v in variables:
            if v.name.startswith("g_") and v.name[2:] not in targets:
                targets.append(v.name[2:])
        # a conductance read but not defined is reset after every step: the
        # variable g_<target> = 0.0 appended after the others
        # (parser/AnalyseNeuron.py:186-209, Extraction.py:686-703)
        for v in list(variables):
            for m in re.finditer(r"(?:^|[^\w.])g_(\w+)", v.eq):
                gname = "g_" + m.group(1)
                if gname not in declared:
                    declared.add(gname)
                    targets.append(m.group(1))
                    variables.append(Variable(name=gname, eq=gname + " = 0.0", locality="local", flags=[],
                                              bounds={}, init=0.0, ctype=precision))
    # random distributions -> rand_k
    randoms = []
    for v in variables:
        transformed = v.eq
        for dist in DISTRIBUTIONS:
            for m in re.finditer(r"(?P<pre>[^\w.]|^)" + dist + r"\(([^()]+)\)", " " + v.eq):
                args = [a.strip() for a in m.group(2).split(",")]
                if len(args) != 2:
                    raise EquationError(f"{dist} requires two arguments: {v.eq}")
                rv = RandomVar(name=f"rand_{len(randoms)}", dist=dist,
                               args=tuple(float(a) for a in args), locality=v.locality)
                randoms.append(rv)
                transformed = transformed.replace(dist + "(" + m.group(2) + ")", rv.name)
        v.eq = transformed
    for tgt in targets:
        for v in variables:
            v.eq = re.sub(r"sum\(\s*" + tgt + r"\s*\)", "__sum_" + tgt + "__", v.eq)
    desc = NeuronDescription(name=name, spiking=spiking, parameters=params, variables=variables,
                             randoms=randoms, targets=targets, refractory=refractory)
    for v in variables:
        if v.name not in declared:
            raise EquationError(f"cannot identify the variable of: {v.eq}")
    tr = _Translator(desc, precision)
    local_names = desc.local_names() + [r.name for r in randoms if r.locality == "local"]
    for v in variables:
        if v.eq == "":
            continue            # attribute without an equation (the implicit r of spiking neurons)
        eq, ite = _extract_ite(v.eq)
        eq = _logical(eq)
        kind = _equation_type(v.name, eq)
        method = "explicit"
        for m in ("exponential", "explicit"):
            if m in v.flags:
                method = m
        for m in ("implicit", "semiimplicit", "midpoint", "rk4", "event-driven"):
            if m in v.flags:
                raise EquationError(f"numerical method '{m}' is not supported by this backend")
        v.method = method
        if kind == "ODE":
            expr = eq
            expr = expr.replace("=", "- (", 1) + ")" if "=" in expr else expr
            expr = expr.replace(" ", "")
            if method == "exponential" and ite is None:
                v.pre_loop, v.cpp, v.switch = tr.exponential(v.name, expr, local_names)
            else:
                v.pre_loop, v.cpp, v.switch = tr.explicit(v.name, expr)
        elif kind == "inc":
            v.cpp = tr.increment(eq)
        elif kind == "simple":
            v.cpp = tr.assignment(eq)
        else:
            raise EquationError(f"not an equation: {v.eq}")
        if ite is not None:
            v.cpp = v.cpp.replace(_ITE_PH, _ite_code(tr, ite))
        # bounds (generator/Utils.py:48-66)
        code = ""
        idx = "%(local_index)s" if v.locality == "local" else "%(global_index)s"
        for b in ("min", "max"):
            if b in v.bounds:
                val = v.bounds[b]
                try:
                    float(val)
                except ValueError:
                    val = tr.expression(val)
                op = "<" if b == "min" else ">"
                code += f"if({v.name}{idx} {op} {val})\n    {v.name}{idx} = {val};\n"
        v.bound_code = code
    if spiking:
        desc.spike_cond = tr.condition(spike)
        for line in _lines(reset):
            kind = "inc" if any(op in line for op in ["+=", "-=", "*=", "/="]) else "simple"
            desc.spike_reset.append(tr.increment(line) if kind == "inc" else tr.assignment(line))
    return desc


def sort_odes(desc: NeuronDescription, locality: str):
    """Group consecutive ODEs into blocks (generator/Utils.py:17-45)."""
    blocks = []
    is_ode = False
    for v in desc.variables:
        if v.cpp == "" or v.locality != locality:
            continue
        if v.switch is not None:
            if is_ode:
                blocks[-1][1].append(v)
            else:
                is_ode = True
                blocks.append(("ode", [v]))
        else:
            if is_ode or not blocks:
                is_ode = False
                blocks.append(("non-ode", [v]))
            else:
                blocks[-1][1].append(v)
    return blocks


def equation_code(desc: NeuronDescription, locality: str, mask: str | None) -> str:
    """Statement list of the update loop (generator/Utils.py:139-265).

    ODE blocks: every increment first, then every switch; switches of
    non-conductance variables are multiplied by the refractory mask
    (Utils.py:69-77); non-ODE statements of non-conductances are wrapped in
    ``if (mask[i]) { ... }``.
    """
    out = []
    for kind, block in sort_odes(desc, locality):
        if kind == "ode":
            for v in block:
                out.append(v.cpp)
            for v in block:
                sw = v.switch
                if mask is not None and not v.name.startswith("g_"):
                    sw = sw.replace(";", "* " + mask + "%(local_index)s;")
                out.append(sw)
                if v.bound_code:
                    out.append(v.bound_code)
        else:
            for v in block:
                code = v.cpp
                if mask is not None and not v.name.startswith("g_"):
                    code = "if (" + mask + "%(local_index)s) { " + code + " }"
                out.append(code)
                if v.bound_code:
                    out.append(v.bound_code)
    return "\n".join(out)


# ---------------------------------------------------------------------------
# Synapse definitions (parser/AnalyseSynapse.py)
# ---------------------------------------------------------------------------

SYN_LOCALITY = {"local": 0, "semiglobal": 1, "global": 2}
_DEP = re.compile(r"(?<![\w.])(pre|post)\.([A-Za-z_]\w*)")


@dataclass
class SynapseDescription:
    """Attributes of a synapse model and the C text of its events and
    continuous updates.

    Locality (AnalyseSynapse.py): ``: projection`` = one value per projection
    ("global"), ``: postsynaptic`` = one value per post-synaptic neuron
    ("semiglobal"), otherwise one value per synapse ("local").  ``w`` is the
    per-synapse weight; it is a variable when an equation defines it.
    Event-driven variables are integrated exactly between events
    (Equation.py:606-649); the other equations run every step in
    update_synapse (explicit or exponential Euler, Equation.py:276-307,
    :559-604).  ``pre.x`` / ``post.x`` read attribute x of the pre / post
    neuron (ParserTemplate.py:14-21).
    """
    name: str
    parameters: list
    variables: list
    spiking: bool = False
    g_code: str | None = None               # transmission of g_target += <expr>; None = w
    pre_code: list = field(default_factory=list)
    post_code: list = field(default_factory=list)
    event_driven: list = field(default_factory=list)   # (name, C text with %(t)s)
    preloop: list = field(default_factory=list)        # (name, C value) hoisted step sizes
    pre_deps: list = field(default_factory=list)
    post_deps: list = field(default_factory=list)
    randoms: list = field(default_factory=list)
    targets: list = field(default_factory=list)

    def attrs(self):
        """(name, locality, init) of every attribute besides w."""
        return [(a.name, a.locality, a.init) for a in self.parameters + self.variables if a.name != "w"]

    def constant_names(self):
        return [a.name for a in self.parameters if "constant" in a.flags]

    def local_names(self):
        return ["w"] + [n for n, loc, _ in self.attrs() if loc == "local"]

    def global_names(self):
        return [n for n, loc, _ in self.attrs() if loc != "local"]

    def has_update(self, locality: str) -> bool:
        return any(v.cpp and v.locality == locality for v in self.variables)


def _deps_to_symbols(text: str, desc: SynapseDescription) -> str:
    def repl(m):
        side, name = m.group(1), m.group(2)
        lst = desc.pre_deps if side == "pre" else desc.post_deps
        if name not in lst:
            lst.append(name)
        return f"_{side}_{name}_"
    return _DEP.sub(repl, text)


def _syn_locality(flags) -> str:
    if "projection" in flags:
        return "global"
    if "postsynaptic" in flags:
        return "semiglobal"
    return "local"


def _bound_code(tr, name, bounds) -> str:
    lhs = tr.ccode(tr.local_dict[name])
    code = ""
    for b in ("min", "max"):
        if b in bounds:
            val = bounds[b]
            try:
                float(val)
            except ValueError:
                val = tr.expression(_DEP.sub(lambda m: f"_{m.group(1)}_{m.group(2)}_", val))
            op = "<" if b == "min" else ">"
            code += f"if({lhs} {op} {val})\n    {lhs} = {val};\n"
    return code


def _wrap_w(name: str, code: str) -> str:
    # updates of w are skipped while learning is disabled (OpenMPGenerator.py:
    # 1137-1139, :1600-1603; Utils.py:105-116 wrap_w)
    return "if (_plasticity) {\n" + code.rstrip("\n") + "\n}" if name == "w" else code


def _statement(tr, line, desc) -> str:
    eq, constraint = _split_constraint(line)
    bounds, _ = _flags(constraint)
    name = _variable_name(eq)
    loc = {n: l for n, l, _ in desc.attrs()}
    if name != "w" and name not in loc:
        raise EquationError(f"unknown synaptic variable in: {line}")
    if name != "w" and loc[name] != "local":
        raise EquationError(f"events can only modify per-synapse variables: {line}")
    code = tr.increment(eq) if any(op in eq for op in ["+=", "-=", "*=", "/="]) else tr.assignment(eq)
    code += "\n" + _bound_code(tr, name, bounds)
    return _wrap_w(name, code)


def analyse_synapse(parameters: str = "", equations: str = "", pre_spike: str | None = None,
                    post_spike: str | None = None, name: str = "synapse",
                    precision: str = "double", constants=None) -> SynapseDescription:
    """Analyse a synapse definition (AnalyseSynapse.py) into the C text of its
    pre- and post-synaptic events and of its continuous updates.
    `constants`: names of the network constants an equation may read."""
    params = []
    for line in _lines(parameters):
        eq, constraint = _split_constraint(line)
        bounds, flags = _flags(constraint)
        if "=" not in eq:
            raise EquationError(f"parameter without value: {line}")
        pname, val = eq.split("=", 1)
        params.append(Variable(name=pname.strip(), eq=eq, locality=_syn_locality(flags), flags=flags,
                               bounds=bounds, init=_parse_value(val), is_parameter=True))
    desc = SynapseDescription(name=name, parameters=params, variables=[],
                              spiking=bool(pre_spike) or bool(post_spike))
    if not desc.spiking and re.search(r"\bt_(pre|post)\b", (equations or "") + (parameters or "")):
        raise EquationError("t_pre / t_post are spike times: only spiking synapses use them")
    for eq, constraint in _equations(equations):
        bounds, flags = _flags(constraint)
        eq = _deps_to_symbols(eq, desc)
        vname = _variable_name(eq)
        init = _parse_value(bounds.pop("init")) if "init" in bounds else 0.0
        method = "explicit"
        for m in ("event-driven", "exponential", "explicit"):
            if m in flags:
                method = m
        for m in ("implicit", "semiimplicit", "midpoint", "rk4"):
            if m in flags:
                raise EquationError(f"numerical method '{m}' is not supported by this backend")
        loc = _syn_locality(flags)
        if vname == "w" and loc != "local":
            raise EquationError("w is a per-synapse variable")
        if method == "event-driven" and (loc != "local" or not desc.spiking):
            raise EquationError(f"event-driven variables are per-synapse variables of spiking synapses: {line}")
        desc.variables.append(Variable(name=vname, eq=eq, locality=loc, flags=flags, bounds=bounds, init=init,
                                       method=method))
    for a in params:
        if a.name == "w":
            raise EquationError("w is the synaptic weight: define it with an equation, not as a parameter")
    declared = {a.name for a in params + desc.variables} | {"w"}
    for c in _constant_params(constants, declared, [equations, pre_spike, post_spike], precision):
        params.append(c)
        declared.add(c.name)
    tr = _Translator(desc, precision)
    # attributes become _s_<name> in the kernels so that no attribute name
    # can collide with a kernel variable
    index = {"local": "%(local_index)s", "semiglobal": "%(semiglobal_index)s", "global": "%(global_index)s"}
    for n, loc, _ in desc.attrs():
        tr.local_dict[n] = sp.Symbol("_s_" + n + index[loc])
    # last spike times of the pre / post neuron in ms (ParserTemplate.py:18-19)
    tr.local_dict["t_pre"] = sp.Symbol("((double)(_pre_last_spike[pre])*dt)")
    tr.local_dict["t_post"] = sp.Symbol("((double)(_post_last_spike[post])*dt)")
    texts = [pre_spike or "", post_spike or ""]
    texts = [_deps_to_symbols(t, desc) for t in texts]
    for side, lst in (("pre", desc.pre_deps), ("post", desc.post_deps)):
        for n in lst:
            idx = "pre" if side == "pre" else "post"
            tr.local_dict[f"_{side}_{n}_"] = sp.Symbol(f"_d{side}_{n}[{idx}*_st_{side}_{n}]")
    local_names = [n for n in desc.local_names()]
    for v in desc.variables:
        if v.method == "event-driven":
            # exact decay since the last event (Equation.py:606-649)
            expr = v.eq.replace("=", "- (", 1) + ")"
            expr = expr.replace(" ", "")
            real_tau, _, steady = tr.standardize(v.name, expr)
            if real_tau is None:
                raise EquationError(f"the equation is not a linear ODE and cannot be evaluated exactly: {v.eq}")
            for other in desc.variables:
                if other.name != v.name and tr.local_dict[other.name] in steady.atoms():
                    raise EquationError(f"an event-driven variable cannot depend on other variables: {v.eq}")
            var = tr.ccode(tr.local_dict[v.name])
            st = tr.ccode(steady)
            decay = f"exp(dt*(_last_event%(local_index)s - %(t)s)/({tr.ccode(real_tau)}))"
            code = f"{var} *= {decay};" if st == "0" else f"{var} = {st} + ({var} - {st})*{decay};"
            desc.event_driven.append((v.name, code))
            v.cpp = ""
            continue
        eq, ite = _extract_ite(v.eq)
        eq = _logical(eq)
        kind = _equation_type(v.name, eq)
        if kind == "ODE":
            expr = eq.replace("=", "- (", 1) + ")"
            expr = expr.replace(" ", "")
            if v.method == "exponential" and ite is None:
                v.pre_loop, v.cpp, v.switch = tr.exponential(v.name, expr, local_names)
            else:
                v.pre_loop, v.cpp, v.switch = tr.explicit(v.name, expr)
            if v.pre_loop:
                desc.preloop.append((v.pre_loop["name"], v.pre_loop["value"]))
        elif kind == "inc":
            v.cpp = tr.increment(eq)
        elif kind == "simple":
            v.cpp = tr.assignment(eq)
        else:
            raise EquationError(f"not an equation: {v.eq}")
        if ite is not None:
            v.cpp = v.cpp.replace(_ITE_PH, _ite_code(tr, ite))
        v.bound_code = _bound_code(tr, v.name, v.bounds)
    for v in desc.variables:
        if v.name not in declared:
            raise EquationError(f"cannot identify the variable of: {v.eq}")
    for line in _lines(texts[0]):
        eq, _ = _split_constraint(line)
        e = eq.replace(" ", "")
        if e.startswith("g_target"):
            if not e.startswith("g_target+="):
                raise EquationError(f"only g_target += <expression> is supported: {line}")
            desc.g_code = tr.expression(eq.split("+=", 1)[1])
            continue
        if "g_target" in eq:
            raise EquationError(f"g_target can only be incremented: {line}")
        desc.pre_code.append(_statement(tr, line, desc))
    for line in _lines(texts[1]):
        desc.post_code.append(_statement(tr, line, desc))
    return desc


def synapse_update_code(desc: SynapseDescription, locality: str) -> str:
    """Statements of update_synapse for one locality (Utils.py:139-265): ODE
    blocks compute every increment, then apply every switch; updates of w
    are wrapped in the learning flag."""
    out = []
    for kind, block in sort_odes(desc, locality):
        if kind == "ode":
            for v in block:
                out.append(v.cpp)
            for v in block:
                out.append(_wrap_w(v.name, v.switch + "\n" + v.bound_code))
        else:
            for v in block:
                out.append(_wrap_w(v.name, v.cpp + "\n" + v.bound_code))
    return "\n".join(out)
