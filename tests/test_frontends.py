"""Right-hand-side front-ends: Python functions (clode_amd/function_converter.py,
counterpart of clode/function_converter.py) and XPP files
(clode_amd/xpp_parser.py, counterpart of clode/xpp_parser.py).

CPU: the reference converter's error cases (test/test_function_converter.py
type-error table), the translation rules, XPP parsing, and hipRTC
compilation of the generated sources (including every OpenCL built-in a
Python right-hand side may call).  GPU: a converted right-hand side and a
converted XPP file integrate bit for bit like the hand-written model file,
and every built-in evaluates on the device like its Python definition
(clode/opencl_builtins.py; test/test_opencl_builtins.py checks it at
atol 1e-7)."""
import math
import os
from typing import List

import numpy as np
import pytest

import clode_amd as clode
from clode_amd import opencl_builtins as ob
from clode_amd.function_converter import OpenCLConverter, convert_str_to_opencl
from clode_amd.xpp_parser import convert_xpp_file, format_opencl_rhs, read_ode_parameters

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(REPO, "clode_amd", "models")


@pytest.mark.parametrize("message, code", [
    ("Unsupported type 'str' at line 1", "def f(x: str) -> float:\n    return x\n"),
    ("Argument 'x' has no type at line 1", "def f(x) -> float:\n    return x\n"),
    ("Cannot change the type of variable 'a'", "def f() -> int:\n    a: int = 1\n    a = 2.2\n    return a\n"),
    ("Function 'g' must have a return type at line 1", "def g():\n    a: int = 1\n    return a\n"),
    ("Unsupported type 'str' at line 1", "def f() -> str:\n    a: int = 1\n    return a\n"),
    ("Unsupported type 'str' at line 2", "def f() -> int:\n    a: str = 2\n    return 1\n"),
    ("Cannot assign multiple variables in one line at line 2", "def f() -> int:\n    a, b = 1, 2\n    return a\n"),
    ("Cannot redeclare variable 'a' at line 3", "def f() -> int:\n    a: int = 1\n    a: int = 2\n    return a\n"),
])
def test_type_errors(message, code):
    with pytest.raises(TypeError, match=message):
        convert_str_to_opencl(code)


def test_unknown_names():
    with pytest.raises(ValueError, match="Variable 'b' not found"):
        convert_str_to_opencl("def f(a: float) -> float:\n    return a + b\n")
    with pytest.raises(ValueError, match="Function 'nope' not found"):
        convert_str_to_opencl("def f(a: float) -> float:\n    return nope(a)\n")


def test_translation_rules():
    c = convert_str_to_opencl(
        "def f(a: float, n: int, v: List[float], out: List[float]) -> None:\n"
        "    x: float = a ** 2 + a ** 3 + a ** 5 + a ** 0.5 + a ** n + a ** 0 + a ** 1\n"
        "    y: float = n / 2 + n // 3 + a % 2.0 + n % 3\n"
        "    out[0] = x if a > 0 and n != 1 else -y\n"
        "    out[1] += math.exp(v[n]) + np.sqrt(abs(a)) + max(a, 1.0) + pi\n")
    assert "const realtype a" in c and "const int n" in c and "const realtype v[]" in c
    assert "realtype out[]" in c and "const realtype out" not in c       # written -> not const
    assert "(a * a)" in c and "(a * a * a)" in c and "pown(a, 5)" in c and "pow(a, RCONST(0.5))" in c
    assert "pown(a, n)" in c
    assert "((realtype)n / 2)" in c and "floor" in c and "fmod(a, RCONST(2.0))" in c and "(n % 3)" in c
    assert "exp(v[n])" in c and "sqrt(fabs(a))" in c and "fmax(a, RCONST(1.0))" in c
    assert "? x : (-y)" in c and "&&" in c


def test_functions_build_on_each_other():
    conv = OpenCLConverter()
    first = conv.convert_source("def sq(a: float) -> float:\n    return a * a\n")
    both = conv.convert_source("def g(x: List[float], d: List[float]) -> None:\n    d[0] = sq(x[0])\n",
                               mutable_args=[1], function_name="getRHS")
    assert both.startswith(first) and "void getRHS(const realtype x[], realtype d[])" in both
    assert "d[0] = sq(x[0]);" in both


XPP = """# Van der Pol
init x = 2.0
init y = 0
par mu = 1.0
% preserved comment
aux x2 = x^2
aux xneg = x^-2
y' = mu * (1 - x*x) * y - x
x' = y
@ dt=0.01
done
"""


def test_xpp_parsing(tmp_path):
    pars, aux, init, dx, noise, st = read_ode_parameters(XPP)
    assert list(init) == ["x", "y"] and list(pars) == ["mu"] and list(aux) == ["x2", "xneg"]
    assert dx["y"] == "mu * (1 - x*x) * y - x" and noise == []
    src = format_opencl_rhs(pars, aux, init, dx, noise, st)
    assert "realtype x = x_[0];" in src and "realtype mu = p_[0];" in src
    assert "realtype x2 = (x * x);" in src and "pown(x, -2)" in src and "/* preserved comment */" in src
    assert src.index("dx_[0] = d_x;") < src.index("dx_[1] = d_y;")  # state order follows init order
    f = tmp_path / "vdp.xpp"
    f.write_text(XPP)
    out = convert_xpp_file(str(f))
    assert out.endswith("vdp.cl") and os.path.exists(out)
    assert read_ode_parameters("par a=1, b = 2*3, c=4")[0] == {"a": "1", "b": "2*3", "c": "4"}


BUILTIN_CALLS = [
    ("acos", "a"), ("acosh", "b"), ("acospi", "a"), ("asin", "a"), ("asinh", "a"), ("asinpi", "a"), ("atan", "a"),
    ("atan2", "a, b"), ("atan2pi", "a, b"), ("atanh", "a"), ("atanpi", "a"), ("cbrt", "c"), ("ceil", "c"),
    ("copysign", "a, m"), ("cos", "c"), ("cosh", "a"), ("cospi", "a"), ("erf", "a"), ("erfc", "a"), ("exp", "c"),
    ("exp2", "c"), ("exp10", "a"), ("expm1", "a"), ("fabs", "m"), ("fdim", "c, a"), ("floor", "c"),
    ("fmod", "c, a"), ("gamma", "c"), ("heaviside", "m"), ("hypot", "a, b"), ("lgamma", "c"), ("log", "c"),
    ("log1p", "a"), ("log2", "c"), ("log10", "c"), ("nextafter", "a, b"), ("pow", "b, c"), ("pown", "b, 3"),
    ("powr", "b, c"), ("remainder", "c, b"), ("rint", "c"), ("rootn", "c, 3"), ("rsqrt", "c"), ("sin", "c"),
    ("sinh", "a"), ("sinpi", "a"), ("sqrt", "c"), ("tan", "a"), ("tanh", "a"), ("tanpi", "a"), ("trunc", "c"),
    ("ldexp", "a, 3"), ("ilogb", "c"),
]
ARGS = {"a": 0.3, "b": 1.7, "c": 2.5, "m": -0.4}


def builtins_rhs_source() -> str:
    lines = ["def rhs(t: float, x: List[float], p: List[float], dx: List[float], aux: List[float],"
             " w: List[float]) -> None:",
             "    a: float = p[0]", "    b: float = p[1]", "    c: float = p[2]", "    m: float = p[3]",
             "    dx[0] = 0.0"]
    for k, (name, args) in enumerate(BUILTIN_CALLS):
        lines.append(f"    aux[{k}] = {name}({args})")
    return "\n".join(lines) + "\n"


def write_builtins_model(tmp_path) -> str:
    src = convert_str_to_opencl(builtins_rhs_source(), mutable_args=[3, 4], function_name="getRHS")
    path = tmp_path / "builtins_rhs.cl"
    path.write_text(src)
    return str(path)


def test_builtins_program_compiles(tmp_path):
    from clode_amd.cpp import clode_cpp_wrapper as w

    path = write_builtins_model(tmp_path)
    aux = [f"f{k}" for k in range(len(BUILTIN_CALLS))]
    for prec in (False, True):
        s = w.TrajectorySimulatorBase(w.ProblemInfo(path, ["x"], ["a", "b", "c", "m"], aux, 0), "euler", prec,
                                      w.OpenCLResource(), "")
        s.compile()


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
def test_builtins_on_device(gpu_available, tmp_path, single):
    path = write_builtins_model(tmp_path)
    aux = [f"f{k}" for k in range(len(BUILTIN_CALLS))]
    sim = clode.TrajectorySimulator(src_file=path, variables={"x": 0.0}, parameters=dict(ARGS), aux=aux,
                                    stepper=clode.Stepper.euler, dt=0.01, dtmax=0.01, t_span=(0.0, 0.0),
                                    max_store=2, single_precision=single)
    traj = sim.trajectory()
    vals = np.array([traj.aux[f"f{k}"][0] for k in range(len(BUILTIN_CALLS))]).ravel()
    for k, (name, args) in enumerate(BUILTIN_CALLS):
        argv = [ARGS[a.strip()] if a.strip() in ARGS else int(a) for a in args.split(",")]
        if single:
            argv = [float(np.float32(v)) if isinstance(v, float) else v for v in argv]
        ref = getattr(ob, name)(*argv)
        tol = 2e-6 if single else 1e-12
        assert abs(vals[k] - ref) <= tol * (1 + abs(ref)), (name, vals[k], ref)


def vdp_python(t: float, x: List[float], p: List[float], dx: List[float], aux: List[float], w: List[float]) -> None:
    mu: float = p[0]
    dx[0] = x[1]
    dx[1] = mu * (1 - x[0] * x[0]) * x[1] - x[0]


@pytest.mark.gpu
def test_converted_sources_match_model_file(gpu_available, tmp_path):
    common = dict(variables={"x": 2.0, "y": 0.0}, parameters={"mu": 1.0}, stepper=clode.Stepper.rk4, dt=0.01,
                  dtmax=0.01, t_span=(0.0, 30.0), single_precision=False)
    mu = np.linspace(0.1, 3.0, 50)
    f = tmp_path / "vdp.xpp"
    f.write_text(XPP)
    results = []
    for kw in (dict(src_file=os.path.join(MODELS, "van_der_pol.cl")), dict(rhs_equation=vdp_python),
               dict(src_file=str(f))):
        sim = clode.Simulator(**kw, **common)
        sim.set_ensemble(parameters={"mu": mu})
        results.append(sim.transient(fetch_results=True))
    assert np.array_equal(results[0], results[1])
    assert np.array_equal(results[0], results[2])
