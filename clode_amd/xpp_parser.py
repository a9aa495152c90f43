"""XPP/XPPAUT ``.ode``/``.xpp`` files -> OpenCL-C right-hand sides.

Counterpart of clode/xpp_parser.py.  The generated text has the layout the
reference produces and tests against a golden file
(test/test_xpp_parser.py:9-22): a ``getRHS`` with one parameter per line and
eight commented sections (state variables, parameters, noise terms, core
equations, auxiliary equations, differential equations, auxiliary outputs,
differential outputs).  Expressions keep their XPP text except for:

- powers ``a^k``: k = 2, 3, 4 become repeated products, other unsigned
  integers ``pown(a, k)``, signed or fractional exponents ``pow(a, k)``;
- decimal literals get an ``f`` suffix (so, as in the reference, they are
  single-precision constants even when ``realtype`` is double);
- ``%`` starts a comment that runs to the end of the line.

Lines understood (at the start of the line, as XPP writes them):

  ``par a=1, b=2`` / ``p ...``     parameters (p_[0], p_[1], ... in order)
  ``init x=0.1, y=0``              state variables (x_[0], ... in order)
  ``x' = expr`` / ``dx/dt = expr`` differential equations
  ``aux name = expr``              auxiliary outputs (aux_[0], ...)
  ``wiener w1 w2``                 noise terms (w_[0], ...)
  ``@ ...``, blank lines           ignored
  ``# ...``, ``done``              ignored (XPP comments and terminator)
  anything else                    copied as a statement (``name = expr``
                                   becomes a ``realtype`` declaration)

A state variable without an equation, or an equation for a variable with no
``init`` line, raises ``ValueError`` (the reference emits code that then
fails to compile).
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

_DEF_NAME = re.compile(r"(\w+)\s*=(?!=)\s*")
_DX_LINE = re.compile(r"^(\w+)'\s*=\s*")
_DXDT_LINE = re.compile(r"^d(\w+)/dt\s*=\s*")
_PREFIXES = {"par": "par", "p": "par", "param": "par", "init": "init", "i": "init", "aux": "aux",
             "wiener": "wiener"}

_POWER = re.compile(r"(\w+)\s*\^\s*([-+]?[\d.]*\d)")
_DECIMAL = re.compile(r"(\d+\.\d*(?:[eE][-+]?\d+)?)")


def _definitions(body: str, line: str) -> List[Tuple[str, str]]:
    """``'a=1, b = 2*c, c=3'`` -> ``[('a', '1'), ('b', '2*c'), ('c', '3')]``.
    A comma ends a value only where another ``name=`` follows it."""
    starts = [m for m in _DEF_NAME.finditer(body)
              if not body[:m.start()].strip() or body[:m.start()].rstrip().endswith(",")]
    if not starts:
        raise ValueError(f"Could not parse line '{line}'")
    out = []
    for k, m in enumerate(starts):
        end = starts[k + 1].start() if k + 1 < len(starts) else len(body)
        out.append((m.group(1), body[m.end():end].strip().rstrip(",").strip()))
    return out


def read_ode_parameters(
    xpp_string: str,
) -> Tuple[Dict[str, str], Dict[str, str], Dict[str, str], Dict[str, str], List[str], List[str]]:
    """(parameters, auxiliaries, initial_values, dx, noise, statements) of an
    XPP file, as in clode/xpp_parser.py:38-75: dicts name -> expression text in
    file order, the noise names, and the remaining lines as statements."""
    parameters: Dict[str, str] = {}
    auxiliaries: Dict[str, str] = {}
    initial_values: Dict[str, str] = {}
    dx: Dict[str, str] = {}
    noise: List[str] = []
    statements: List[str] = []
    stores = {"par": parameters, "init": initial_values, "aux": auxiliaries}
    for line in xpp_string.splitlines():
        head, _, rest = line.partition(" ")
        kind = _PREFIXES.get(head) if rest else None
        if kind == "wiener":
            noise.extend(rest.replace(",", " ").split())
        elif kind is not None:
            stores[kind].update(_definitions(rest, line))
        elif _DX_LINE.match(line) or _DXDT_LINE.match(line):
            m = _DX_LINE.match(line) or _DXDT_LINE.match(line)
            dx[m.group(1)] = line[m.end():]
        elif not line.strip() or line.startswith("@") or line.startswith("#") or line.strip() == "done":
            continue
        else:
            statements.append(line)
    return parameters, auxiliaries, initial_values, dx, noise, statements


def _power(m: re.Match) -> str:
    base, exponent = m.group(1), m.group(2)
    if exponent in ("2", "3", "4"):
        return "*".join([base] * int(exponent))
    if exponent.isdigit():
        return f"pown({base}, {exponent})"
    return f"pow({base}, {exponent})"


def _rewrite_line(line: str) -> str:
    line = _DECIMAL.sub(r"\1f", _POWER.sub(_power, line))
    cut = line.find("%")
    return line if cut < 0 else f"{line[:cut]}/* {line[cut + 1:]} */"


def format_opencl_rhs(parameters, auxiliaries, initial_values, dx, noise, statements) -> str:
    """OpenCL-C ``getRHS`` text for the parsed file (layout of
    clode/xpp_parser.py:78-177)."""
    variables = list(initial_values)
    missing = [v for v in variables if v not in dx]
    if missing:
        raise ValueError(f"No differential equation for {', '.join(missing)}")
    unknown = [v for v in dx if v not in initial_values]
    if unknown:
        raise ValueError(f"No init line for {', '.join(unknown)}")

    def declare(line: str) -> str:
        if "=" not in line:
            return f"    {line};"
        indent = len(line) - len(line.lstrip())
        return f"    {line[:indent]}realtype {line[indent:]};"

    sections = [
        ("State variables", [f"    realtype {v} = x_[{k}];" for k, v in enumerate(variables)]),
        ("Parameters", [f"    realtype {p} = p_[{k}];" for k, p in enumerate(parameters)]),
        ("Noise terms", [f"    realtype {w} = w_[{k}];" for k, w in enumerate(noise)]),
        ("Core equations", [declare(s) for s in statements]),
        ("Auxiliary equations", [f"    realtype {a} = {e};" for a, e in auxiliaries.items()]),
        ("Differential equations", [f"    realtype d{v} = {e};" for v, e in dx.items()]),
        ("Auxiliary outputs", [f"    aux_[{k}] = {a};" for k, a in enumerate(auxiliaries)]),
        ("Differential outputs", [f"    dx_[{k}] = d{v};" for k, v in enumerate(variables)]),
    ]
    params = ["const realtype t", "const realtype x_[]", "const realtype p_[]", "realtype dx_[]",
              "realtype aux_[]", "const realtype w_[]"]
    lead = "void getRHS("
    out = [lead + (",\n" + " " * len(lead)).join(params) + ") {", ""]
    for k, (title, body) in enumerate(sections):
        out.append(f"    /* {title} */")
        out.extend(body)
        if k + 1 < len(sections):
            out.append("")
    out.append("}")
    return "\n".join(_rewrite_line(line) for line in "\n".join(out).split("\n"))


def convert_xpp_file(filename: str) -> str:
    """Write ``<name>.cl`` beside ``<name>.xpp`` / ``<name>.ode`` and return its path."""
    with open(filename) as f:
        text = f.read()
    out = filename[:-4] + ".cl"
    with open(out, "w") as f:
        f.write(format_opencl_rhs(*read_ode_parameters(text)))
    return out
