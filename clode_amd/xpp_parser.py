"""XPP/XPPAUT ``.ode`` files -> OpenCL-C right-hand sides.

Counterpart of clode/xpp_parser.py.  Understood lines:

  ``par a=1, b=2`` / ``p ...``    parameters (in order: p_[0], p_[1], ...)
  ``init x=0.1, y=0``              state variables and initial values (x_[0], ...)
  ``x' = expr`` / ``dx/dt = expr`` differential equations
  ``aux name = expr``              auxiliary outputs (aux_[0], ...)
  ``wiener w``                     noise terms (w_[0], ...)
  ``name = expr``                  intermediate quantities, in file order
  ``% ...``                        comments, kept as C comments
  ``# ...``, ``@ ...``, ``done``   ignored

Expressions use XPP syntax; ``^`` and ``**`` powers become multiplication
(exponents 2..4), ``pown`` (other integers) or ``pow``, and float literals
become ``RCONST`` literals, through the expression translator of
clode_amd/function_converter.py.  State variables that only appear in
``x' =`` lines (without ``init``) are added in order of appearance.
"""
from __future__ import annotations

import ast
import re
from typing import Dict, List, Tuple

from .function_converter import REAL, REAL_ARRAY, _Function

_DEF = re.compile(r"(\w+)\s*=\s*")


def _definitions(body: str) -> List[Tuple[str, str]]:
    """'a=1, b = 2*c, c=3' -> [(a, 1), (b, 2*c), (c, 3)]; a comma separates
    definitions only where a new 'name=' follows."""
    out: List[Tuple[str, str]] = []
    matches = [m for m in _DEF.finditer(body) if body[m.end():m.end() + 1] != "="
               and (m.start() == 0 or body[:m.start()].rstrip().endswith(",") or not body[:m.start()].strip())]
    for k, m in enumerate(matches):
        end = matches[k + 1].start() if k + 1 < len(matches) else len(body)
        value = body[m.end():end].strip().rstrip(",").strip()
        out.append((m.group(1), value))
    if not out and body.strip():
        raise ValueError(f"Could not parse definitions '{body.strip()}'")
    return out


def read_ode_parameters(xpp_string: str):
    """(parameters, auxiliaries, initial_values, dx, noise, statements) of an
    XPP file: dicts name -> expression text (in file order), the noise names,
    and the intermediate statements / comments as (name, expr) or ('%', text)."""
    parameters: Dict[str, str] = {}
    auxiliaries: Dict[str, str] = {}
    initial_values: Dict[str, str] = {}
    dx: Dict[str, str] = {}
    noise: List[str] = []
    statements: List[Tuple[str, str]] = []
    for raw in xpp_string.splitlines():
        line = raw.strip()
        if not line or line.startswith("#") or line.startswith("@") or line.lower() == "done":
            continue
        if line.startswith("%"):
            statements.append(("%", line[1:].strip()))
            continue
        head, _, rest = line.partition(" ")
        if head in ("par", "p", "param"):
            parameters.update(_definitions(rest))
        elif head in ("init", "i"):
            initial_values.update(_definitions(rest))
        elif head == "aux":
            auxiliaries.update(_definitions(rest))
        elif head == "wiener":
            noise.extend(n.strip() for n in rest.replace(",", " ").split() if n.strip())
        else:
            m = re.match(r"^(\w+)'\s*=\s*(.+)$", line) or re.match(r"^d(\w+)/dt\s*=\s*(.+)$", line)
            if m:
                dx[m.group(1)] = m.group(2).strip()
                continue
            m = re.match(r"^(\w+)\s*=\s*(.+)$", line)
            if m:
                statements.append((m.group(1), m.group(2).strip()))
                continue
            raise ValueError(f"Could not parse line '{raw}'")
    return parameters, auxiliaries, initial_values, dx, noise, statements


def _translate(expr: str, fn: _Function) -> str:
    py = expr.replace("^", "**")
    return fn.expr(ast.parse(py, mode="eval").body).text


def format_opencl_rhs(parameters, auxiliaries, initial_values, dx, noise, statements) -> str:
    variables = list(initial_values) + [v for v in dx if v not in initial_values]
    missing = [v for v in variables if v not in dx]
    if missing:
        raise ValueError(f"No differential equation for {', '.join(missing)}")
    stub = ast.parse("def getRHS(t: float) -> None:\n    pass\n").body[0]
    fn = _Function(stub, {}, (), "getRHS")
    scope = fn.scopes[0]
    lines = ["void getRHS(const realtype t, const realtype x_[], const realtype p_[],",
             "            realtype dx_[], realtype aux_[], const realtype w_[])", "{"]
    for k, v in enumerate(variables):
        lines.append(f"    realtype {v} = x_[{k}];")
        scope[v] = REAL
    for k, p in enumerate(parameters):
        lines.append(f"    realtype {p} = p_[{k}];")
        scope[p] = REAL
    for k, w in enumerate(noise):
        lines.append(f"    realtype {w} = w_[{k}];")
        scope[w] = REAL
    for name, value in statements:
        if name == "%":
            lines.append(f"    /* {value} */")
            continue
        lines.append(f"    realtype {name} = {_translate(value, fn)};")
        scope[name] = REAL
    for name, value in auxiliaries.items():
        lines.append(f"    realtype {name} = {_translate(value, fn)};")
        scope[name] = REAL
    for v in variables:
        lines.append(f"    realtype d_{v} = {_translate(dx[v], fn)};")
    for k, name in enumerate(auxiliaries):
        lines.append(f"    aux_[{k}] = {name};")
    for k, v in enumerate(variables):
        lines.append(f"    dx_[{k}] = d_{v};")
    lines.append("}")
    return "\n".join(lines) + "\n"


def convert_xpp_file(filename: str) -> str:
    """Write ``<name>.cl`` beside ``<name>.xpp`` / ``<name>.ode`` and return its path."""
    with open(filename) as f:
        text = f.read()
    out = filename[:-4] + ".cl"
    with open(out, "w") as f:
        f.write(format_opencl_rhs(*read_ode_parameters(text)))
    return out
