"""Python functions -> OpenCL-C right-hand sides for the hipRTC program.

Counterpart of clode/function_converter.py: a typed subset of Python is
translated into the C dialect that the device program compiles (realtype,
RCONST and the OpenCL built-ins of clode_amd/csrc/device/clode_builtins.h).
The converter is written from the language it accepts:

  * functions with annotated arguments (float, int, bool, List[float],
    List[int]) and an annotated return type (None for void);
  * statements: annotated declarations (``a: float = ...``), assignments to
    declared variables and array elements, augmented assignments,
    ``return``, ``if`` / ``elif`` / ``else`` (a variable declared inside a
    branch is local to it), docstrings and ``pass``;
  * expressions: int and float literals, variables, ``x[i]``, the
    arithmetic operators (``**`` becomes repeated multiplication for
    exponents 2..4, ``pown`` for other integers and ``pow`` otherwise;
    ``/`` between ints divides as reals; ``%`` on reals is ``fmod``),
    comparisons, ``and`` / ``or`` / ``not``, conditional expressions,
    calls of OpenCL built-ins (bare or as ``math.`` / ``np.`` / ``numpy.``
    attributes), of ``int`` / ``float`` and of functions converted earlier,
    and ``pi``.

Errors follow the reference converter's conventions: type problems raise
TypeError ("Unsupported type 'str' at line 1", "Argument 'x' has no type at
line 1", "Function 'f' must have a return type at line 1", "Cannot change
the type of variable 'a'", "Cannot redeclare variable 'a' at line 3",
"Cannot assign multiple variables in one line at line 2"), unknown names
raise ValueError.  Line numbers count from the function's ``def``.
"""
from __future__ import annotations

import ast
import inspect
import textwrap
from typing import Callable, Dict, List, Optional, Sequence

OpenCLRhsEquation = Callable[..., None]

REAL, INT, BOOL, VOID = "realtype", "int", "bool", "void"
REAL_ARRAY, INT_ARRAY = "realtype[]", "int[]"

# name -> (C name, argument count, return type) (clode/function_converter.py:55-125)
_BUILTINS: Dict[str, tuple] = {}
for _n in ("acos", "acosh", "acospi", "asin", "asinh", "asinpi", "atan", "atanh", "atanpi", "cbrt", "ceil", "cos",
           "cosh", "cospi", "erf", "erfc", "exp", "exp2", "exp10", "expm1", "fabs", "floor", "lgamma", "log",
           "log2", "log10", "log1p", "rint", "round", "rsqrt", "sin", "sinh", "sinpi", "sqrt", "tan", "tanh",
           "tanpi", "trunc", "heaviside"):
    _BUILTINS[_n] = (_n, 1, REAL)
for _n in ("atan2", "atan2pi", "copysign", "fdim", "fmod", "hypot", "nextafter", "pow", "powr", "remainder",
           "ldexp", "fmax", "fmin"):
    _BUILTINS[_n] = (_n, 2, REAL)
_BUILTINS.update({
    "abs": ("fabs", 1, REAL),
    "max": ("fmax", 2, REAL),
    "min": ("fmin", 2, REAL),
    "mod": ("fmod", 2, REAL),
    "gamma": ("tgamma", 1, REAL),
    "pown": ("pown", 2, REAL),
    "rootn": ("rootn", 2, REAL),
    "ilogb": ("ilogb", 1, INT),
})
_MODULES = ("math", "np", "numpy", "cl", "opencl_builtins")
_CONSTANTS = {"pi": "RCONST(3.141592653589793)", "e": "RCONST(2.718281828459045)"}

_BINOPS = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Mod: "%", ast.FloorDiv: "//",
           ast.Pow: "**"}
_CMPOPS = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">=", ast.Eq: "==", ast.NotEq: "!="}


def _scalar(t: str) -> bool:
    return t in (REAL, INT, BOOL)


class _Expr:
    __slots__ = ("text", "type")

    def __init__(self, text: str, typ: str):
        self.text, self.type = text, typ


def _annotation_type(node: Optional[ast.expr], line: int) -> str:
    """float, int, bool, None, List[float] / list[float] / Sequence[float], List[int]."""
    if node is None:
        raise TypeError(f"Missing type at line {line}")
    if isinstance(node, ast.Constant) and node.value is None:
        return VOID
    if isinstance(node, ast.Name):
        simple = {"float": REAL, "int": INT, "bool": BOOL, "None": VOID}
        if node.id in simple:
            return simple[node.id]
        raise TypeError(f"Unsupported type '{node.id}' at line {line}")
    if isinstance(node, ast.Attribute):
        raise TypeError(f"Unsupported type '{node.attr}' at line {line}")
    if isinstance(node, ast.Subscript):
        outer = node.value.id if isinstance(node.value, ast.Name) else getattr(node.value, "attr", "?")
        if outer in ("List", "list", "Sequence", "MutableSequence"):
            inner = _annotation_type(node.slice, line)
            if inner == REAL:
                return REAL_ARRAY
            if inner == INT:
                return INT_ARRAY
            raise TypeError(f"Unsupported element type for '{outer}' at line {line}")
        raise TypeError(f"Unsupported type '{outer}' at line {line}")
    if isinstance(node, ast.Constant) and isinstance(node.value, str):
        return _annotation_type(ast.parse(node.value, mode="eval").body, line)
    raise TypeError(f"Unsupported type at line {line}")


def _assignable(target: str, value: str) -> bool:
    if target == value:
        return True
    return (target == REAL and value in (INT, BOOL)) or (target == INT and value == BOOL) or \
           (target == BOOL and value == INT)


class _Function:
    """Translation of one FunctionDef."""

    def __init__(self, node: ast.FunctionDef, known: Dict[str, str], mutable_args: Sequence[int],
                 name: Optional[str]):
        self.node = node
        self.known = known
        self.name = name or node.name
        self.line0 = node.lineno - 1
        self.scopes: List[Dict[str, str]] = [{}]
        self.lines: List[str] = []
        if node.returns is None:
            raise TypeError(f"Function '{node.name}' must have a return type at line {self.line(node)}")
        self.ret = _annotation_type(node.returns, self.line(node))
        written = self._written_names(node)
        self.args = []
        for k, a in enumerate(node.args.args):
            if a.annotation is None:
                raise TypeError(f"Argument '{a.arg}' has no type at line {self.line(node)}")
            t = _annotation_type(a.annotation, self.line(node))
            if t == VOID:
                raise TypeError(f"Argument '{a.arg}' cannot be None at line {self.line(node)}")
            const = not (k in mutable_args or a.arg in written)
            self.args.append((a.arg, t, const))
            self.scopes[0][a.arg] = t

    def line(self, node: ast.AST) -> int:
        return getattr(node, "lineno", self.line0 + 1) - self.line0

    @staticmethod
    def _written_names(node: ast.FunctionDef) -> set:
        out = set()
        for n in ast.walk(node):
            targets = []
            if isinstance(n, (ast.Assign,)):
                targets = n.targets
            elif isinstance(n, (ast.AugAssign, ast.AnnAssign)):
                targets = [n.target]
            for t in targets:
                if isinstance(t, ast.Subscript) and isinstance(t.value, ast.Name):
                    out.add(t.value.id)
                elif isinstance(t, ast.Name):
                    out.add(t.id)
        return out

    # ---- names
    def lookup(self, name: str) -> Optional[str]:
        for s in reversed(self.scopes):
            if name in s:
                return s[name]
        return None

    # ---- expressions
    def expr(self, e: ast.expr) -> _Expr:
        line = self.line(e)
        if isinstance(e, ast.Constant):
            if isinstance(e.value, bool):
                return _Expr("1" if e.value else "0", BOOL)
            if isinstance(e.value, int):
                return _Expr(str(e.value), INT)
            if isinstance(e.value, float):
                return _Expr(f"RCONST({e.value!r})", REAL)
            raise ValueError(f"Unsupported constant '{e.value}' at line {line}")
        if isinstance(e, ast.Name):
            t = self.lookup(e.id)
            if t is None:
                if e.id in _CONSTANTS:
                    return _Expr(_CONSTANTS[e.id], REAL)
                raise ValueError(f"Variable '{e.id}' not found at line {line}")
            return _Expr(e.id, t)
        if isinstance(e, ast.Attribute):
            if isinstance(e.value, ast.Name) and e.value.id in _MODULES and e.attr in _CONSTANTS:
                return _Expr(_CONSTANTS[e.attr], REAL)
            raise ValueError(f"Unsupported attribute '{e.attr}' at line {line}")
        if isinstance(e, ast.Subscript):
            return self.subscript(e)
        if isinstance(e, ast.UnaryOp):
            a = self.expr(e.operand)
            if isinstance(e.op, ast.USub):
                return _Expr(f"(-{a.text})", a.type)
            if isinstance(e.op, ast.UAdd):
                return a
            if isinstance(e.op, ast.Not):
                return _Expr(f"(!{a.text})", BOOL)
            raise ValueError(f"Unsupported unary operator at line {line}")
        if isinstance(e, ast.BinOp):
            return self.binop(e)
        if isinstance(e, ast.Compare):
            parts, left = [], self.expr(e.left)
            for op, right_node in zip(e.ops, e.comparators):
                if type(op) not in _CMPOPS:
                    raise ValueError(f"Unsupported comparison at line {line}")
                right = self.expr(right_node)
                parts.append(f"({left.text} {_CMPOPS[type(op)]} {right.text})")
                left = right
            return _Expr(parts[0] if len(parts) == 1 else "(" + " && ".join(parts) + ")", BOOL)
        if isinstance(e, ast.BoolOp):
            op = " && " if isinstance(e.op, ast.And) else " || "
            return _Expr("(" + op.join(self.expr(v).text for v in e.values) + ")", BOOL)
        if isinstance(e, ast.IfExp):
            c, a, b = self.expr(e.test), self.expr(e.body), self.expr(e.orelse)
            t = a.type if a.type == b.type else REAL
            return _Expr(f"({c.text} ? {a.text} : {b.text})", t)
        if isinstance(e, ast.Call):
            return self.call(e)
        raise ValueError(f"Unsupported expression '{type(e).__name__}' at line {line}")

    def subscript(self, e: ast.Subscript) -> _Expr:
        line = self.line(e)
        if not isinstance(e.value, ast.Name):
            raise ValueError(f"Unsupported array expression at line {line}")
        t = self.lookup(e.value.id)
        if t is None:
            raise ValueError(f"Variable '{e.value.id}' not found at line {line}")
        if t not in (REAL_ARRAY, INT_ARRAY):
            raise TypeError(f"Variable '{e.value.id}' is not an array at line {line}")
        idx = self.expr(e.slice)
        if idx.type not in (INT, BOOL):
            raise TypeError(f"Array index must be an integer at line {line}")
        return _Expr(f"{e.value.id}[{idx.text}]", REAL if t == REAL_ARRAY else INT)

    def binop(self, e: ast.BinOp) -> _Expr:
        line = self.line(e)
        op = _BINOPS.get(type(e.op))
        if op is None:
            raise ValueError(f"Unsupported operator at line {line}")
        a = self.expr(e.left)
        if op == "**":
            n = e.right
            neg = isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub)
            lit = n.operand if neg else n
            if isinstance(lit, ast.Constant) and isinstance(lit.value, int) and not isinstance(lit.value, bool):
                k = -lit.value if neg else lit.value
                if k == 0:
                    return _Expr("1", INT if a.type == INT else REAL)
                if k == 1:
                    return a
                if 2 <= k <= 4:
                    return _Expr("(" + " * ".join([a.text] * k) + ")", a.type)
                return _Expr(f"pown({a.text}, {k})", REAL)
            b = self.expr(n)
            if b.type == INT:
                return _Expr(f"pown({a.text}, {b.text})", REAL)
            return _Expr(f"pow({a.text}, {b.text})", REAL)
        b = self.expr(e.right)
        t = REAL if REAL in (a.type, b.type) else INT
        if op == "/":
            if t == INT:
                return _Expr(f"((realtype){a.text} / {b.text})", REAL)
            return _Expr(f"({a.text} / {b.text})", REAL)
        if op == "//":
            if t == INT:
                return _Expr(f"((int)floor((realtype){a.text} / {b.text}))", INT)
            return _Expr(f"floor({a.text} / {b.text})", REAL)
        if op == "%":
            if t == REAL:
                return _Expr(f"fmod({a.text}, {b.text})", REAL)
            return _Expr(f"({a.text} % {b.text})", INT)
        return _Expr(f"({a.text} {op} {b.text})", t)

    def call(self, e: ast.Call) -> _Expr:
        line = self.line(e)
        if e.keywords:
            raise ValueError(f"Keyword arguments are not supported at line {line}")
        if isinstance(e.func, ast.Name):
            fname = e.func.id
        elif isinstance(e.func, ast.Attribute) and isinstance(e.func.value, ast.Name):
            fname = e.func.attr
            if e.func.value.id not in _MODULES:
                raise ValueError(f"Function '{fname}' from module '{e.func.value.id}' not found at line {line}")
        else:
            raise ValueError(f"Unsupported function call at line {line}")
        args = [self.expr(a) for a in e.args]
        if fname in ("int", "float") and isinstance(e.func, ast.Name):
            if len(args) != 1:
                raise ValueError(f"'{fname}' takes one argument at line {line}")
            t = INT if fname == "int" else REAL
            return _Expr(f"(({t})({args[0].text}))", t)
        if fname in self.known and isinstance(e.func, ast.Name):
            return _Expr(f"{fname}({', '.join(a.text for a in args)})", self.known[fname])
        if fname in _BUILTINS:
            cname, nargs, rt = _BUILTINS[fname]
            if len(args) != nargs:
                raise ValueError(f"'{fname}' takes {nargs} argument(s) at line {line}")
            return _Expr(f"{cname}({', '.join(a.text for a in args)})", rt)
        raise ValueError(f"Function '{fname}' not found at line {line}")

    # ---- statements
    def emit(self, depth: int, text: str) -> None:
        self.lines.append("    " * depth + text)

    def body(self, stmts: List[ast.stmt], depth: int) -> None:
        for s in stmts:
            self.stmt(s, depth)

    def stmt(self, s: ast.stmt, depth: int) -> None:
        line = self.line(s)
        if isinstance(s, ast.Expr) and isinstance(s.value, ast.Constant) and isinstance(s.value.value, str):
            return  # docstring
        if isinstance(s, ast.Pass):
            return
        if isinstance(s, ast.AnnAssign):
            if not isinstance(s.target, ast.Name):
                raise TypeError(f"Only variables can be declared at line {line}")
            name = s.target.id
            if name in self.scopes[-1]:
                raise TypeError(f"Cannot redeclare variable '{name}' at line {line}")
            t = _annotation_type(s.annotation, line)
            if not _scalar(t):
                raise TypeError(f"Unsupported type for variable '{name}' at line {line}")
            if s.value is None:
                self.scopes[-1][name] = t
                self.emit(depth, f"{t} {name};")
                return
            v = self.expr(s.value)
            if not _assignable(t, v.type):
                raise TypeError(f"Cannot change the type of variable '{name}' at line {line}")
            self.scopes[-1][name] = t
            self.emit(depth, f"{t} {name} = {v.text};")
            return
        if isinstance(s, ast.Assign):
            if len(s.targets) != 1 or isinstance(s.targets[0], ast.Tuple):
                raise TypeError(f"Cannot assign multiple variables in one line at line {line}")
            tgt = s.targets[0]
            v = self.expr(s.value)
            if isinstance(tgt, ast.Name):
                t = self.lookup(tgt.id)
                if t is None:
                    raise ValueError(f"Variable '{tgt.id}' not declared at line {line}")
                if not _assignable(t, v.type):
                    raise TypeError(f"Cannot change the type of variable '{tgt.id}' at line {line}")
                self.emit(depth, f"{tgt.id} = {v.text};")
                return
            if isinstance(tgt, ast.Subscript):
                ref = self.subscript(tgt)
                if not _assignable(ref.type, v.type):
                    raise TypeError(f"Cannot change the type of an element of '{tgt.value.id}' at line {line}")
                self.emit(depth, f"{ref.text} = {v.text};")
                return
            raise ValueError(f"Variable not declared at line {line}")
        if isinstance(s, ast.AugAssign):
            op = _BINOPS.get(type(s.op))
            if op not in ("+", "-", "*", "/"):
                raise ValueError(f"Unsupported augmented assignment at line {line}")
            ref = self.subscript(s.target) if isinstance(s.target, ast.Subscript) else self.expr(s.target)
            v = self.expr(s.value)
            if ref.type == INT and (v.type == REAL or op == "/"):
                raise TypeError(f"Cannot change the type of variable '{ref.text}' at line {line}")
            self.emit(depth, f"{ref.text} {op}= {v.text};")
            return
        if isinstance(s, ast.Return):
            if s.value is None:
                if self.ret != VOID:
                    raise TypeError(f"Function '{self.node.name}' must return a value at line {line}")
                self.emit(depth, "return;")
                return
            if self.ret == VOID:
                raise TypeError(f"Function '{self.node.name}' returns None at line {line}")
            v = self.expr(s.value)
            if not _assignable(self.ret, v.type):
                raise TypeError(f"Return type mismatch in '{self.node.name}' at line {line}")
            self.emit(depth, f"return {v.text};")
            return
        if isinstance(s, ast.If):
            c = self.expr(s.test)
            self.emit(depth, f"if ({c.text}) {{")
            self.scopes.append({})
            self.body(s.body, depth + 1)
            self.scopes.pop()
            if s.orelse:
                self.emit(depth, "} else {")
                self.scopes.append({})
                self.body(s.orelse, depth + 1)
                self.scopes.pop()
            self.emit(depth, "}")
            return
        raise TypeError(f"Unsupported statement '{type(s).__name__}' at line {line}")

    def translate(self) -> str:
        args = []
        for name, t, const in self.args:
            base = t[:-2] if t.endswith("[]") else t
            suffix = "[]" if t.endswith("[]") else ""
            args.append(f"{'const ' if const else ''}{base} {name}{suffix}")
        self.body(self.node.body, 1)
        # one parameter per line, aligned under the first (the layout the
        # reference's converter tests expect, test_function_converter.py:85-181)
        lead = f"{self.ret} {self.name}("
        head = lead + (",\n" + " " * len(lead)).join(args) + ") {"
        return "\n".join([head] + self.lines + ["}"]) + "\n\n"


class OpenCLConverter:
    """Accumulates converted functions; each conversion returns the program
    text of every function converted so far, in order (later functions may
    call earlier ones)."""

    def __init__(self, entry_function_name: str = "getRHS"):
        self.entry_function_name = entry_function_name
        self._functions: List[str] = []
        self._known: Dict[str, str] = {}

    def _convert_node(self, node: ast.FunctionDef, mutable_args: Sequence[int], function_name: Optional[str]) -> str:
        fn = _Function(node, self._known, mutable_args or (), function_name)
        text = fn.translate()
        self._functions.append(text)
        self._known[fn.name] = fn.ret
        return "".join(self._functions)

    def convert_to_opencl(self, function: Callable, mutable_args: Optional[Sequence[int]] = None,
                          function_name: Optional[str] = None) -> str:
        src = textwrap.dedent(inspect.getsource(function))
        tree = ast.parse(src)
        defs = [n for n in tree.body if isinstance(n, ast.FunctionDef)]
        if not defs:
            raise ValueError("no function definition found")
        return self._convert_node(defs[0], mutable_args or (), function_name)

    def convert_source(self, source: str, mutable_args: Optional[Sequence[int]] = None,
                       function_name: Optional[str] = None) -> str:
        tree = ast.parse(textwrap.dedent(source))
        defs = [n for n in tree.body if isinstance(n, ast.FunctionDef)]
        if not defs:
            raise ValueError("no function definition found")
        out = ""
        for k, node in enumerate(defs):
            last = k == len(defs) - 1
            out = self._convert_node(node, mutable_args if last else (), function_name if last else None)
        return out


def convert_str_to_opencl(source: str, mutable_args: Optional[Sequence[int]] = None,
                          function_name: Optional[str] = None) -> str:
    """Convert every function in ``source`` (the last one may be renamed and
    given mutable array arguments)."""
    return OpenCLConverter().convert_source(source, mutable_args, function_name)
