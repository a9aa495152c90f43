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
from clode_amd.opencl_builtins import *  # noqa: F401,F403  (names used by _every_builtin)
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


def _add_float(a: float, b: float) -> float:
    res: float = a + b
    return res


def _add_floats_in_list(lst_in: List[float], lst_out: List[float]) -> None:
    res: float = lst_in[0] + lst_in[1]
    lst_out[0] = res


def _all_operations(a: float, b: float, c: int, d: int) -> float:
    res1: float = a + b * c / d
    res2: float = ((a - b) % c) ** d
    res3: float = a**b
    res4: float = a**1
    res5: float = a**2
    res6: float = a**3
    res7: float = a**4
    res8: float = a**5
    res9: float = a**0
    res10: float = a**0.5
    sum_res: float = (
        res1 + res2 + -res3 + res4 + res5 + res6 + res7 + res8 + res9 + res10
    )
    return res1 + sum_res


def _get_rhs(var: List[float], derivatives: List[float]) -> None:
    res: float = _add_float(var[0], var[1])
    derivatives[0] = res


# Expected texts of test/test_function_converter.py:79-181, with the function
# names of this module.  One deliberate difference: a floating-point ``%``
# becomes ``fmod`` (the reference emits ``%``, which OpenCL C rejects for
# floating-point operands).
CONVERSIONS = [
    (_add_float,
     "realtype _add_float(const realtype a,\n"
     "                    const realtype b) {\n"
     "    realtype res = (a + b);\n"
     "    return res;\n"
     "}\n\n"),
    (_add_floats_in_list,
     "void _add_floats_in_list(const realtype lst_in[],\n"
     "                         realtype lst_out[]) {\n"
     "    realtype res = (lst_in[0] + lst_in[1]);\n"
     "    lst_out[0] = res;\n"
     "}\n\n"),
    (_all_operations,
     "realtype _all_operations(const realtype a,\n"
     "                         const realtype b,\n"
     "                         const int c,\n"
     "                         const int d) {\n"
     "    realtype res1 = (a + ((b * c) / d));\n"
     "    realtype res2 = pown(fmod((a - b), c), d);\n"
     "    realtype res3 = pow(a, b);\n"
     "    realtype res4 = a;\n"
     "    realtype res5 = (a * a);\n"
     "    realtype res6 = (a * a * a);\n"
     "    realtype res7 = (a * a * a * a);\n"
     "    realtype res8 = pown(a, 5);\n"
     "    realtype res9 = 1;\n"
     "    realtype res10 = pow(a, RCONST(0.5));\n"
     "    realtype sum_res = (((((((((res1 + res2) + (-res3)) + res4) + res5) + res6) + res7) + res8) + res9)"
     " + res10);\n"
     "    return (res1 + sum_res);\n"
     "}\n\n"),
]


@pytest.mark.parametrize("fn, expected", CONVERSIONS, ids=lambda v: getattr(v, "__name__", ""))
def test_reference_conversions(fn, expected):
    assert OpenCLConverter().convert_to_opencl(fn) == expected


def test_reference_second_function_referencing_first():
    conv = OpenCLConverter()
    conv.convert_to_opencl(_add_float)
    assert conv.convert_to_opencl(_get_rhs) == CONVERSIONS[0][1] + (
        "void _get_rhs(const realtype var[],\n"
        "              realtype derivatives[]) {\n"
        "    realtype res = _add_float(var[0], var[1]);\n"
        "    derivatives[0] = res;\n"
        "}\n\n")


def test_functions_build_on_each_other():
    conv = OpenCLConverter()
    first = conv.convert_source("def sq(a: float) -> float:\n    return a * a\n")
    both = conv.convert_source("def g(x: List[float], d: List[float]) -> None:\n    d[0] = sq(x[0])\n",
                               mutable_args=[1], function_name="getRHS")
    assert both.startswith(first) and "void getRHS(const realtype x[],\n            realtype d[]) {" in both
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
    assert "realtype x2 = x*x;" in src and "pow(x, -2)" in src and "/*  preserved comment; */" in src
    assert "realtype dy = mu * (1 - x*x) * y - x;" in src and "@" not in src and "done" not in src
    assert src.index("dx_[0] = dx;") < src.index("dx_[1] = dy;")  # state order follows init order
    assert "pown(a, 7)" in format_opencl_rhs({}, {}, {"a": "1"}, {"a": "a^7 + 2.5e-3"}, [], [])
    assert "2.5e-3f" in format_opencl_rhs({}, {}, {"a": "1"}, {"a": "a^7 + 2.5e-3"}, [], [])
    with pytest.raises(ValueError, match="No differential equation"):
        format_opencl_rhs({}, {}, {"a": "1", "b": "0"}, {"a": "a"}, [], [])
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


GOLDEN_XPP = os.path.join(REPO, "tests", "golden", "xpp")


def test_xpp_golden_file(tmp_path):
    """test/test_xpp_parser.py:9-19: the converted file equals the reference's
    golden output byte for byte (tests/golden/xpp holds both files)."""
    src = tmp_path / "van_der_pol_oscillator.xpp"
    with open(os.path.join(GOLDEN_XPP, "van_der_pol_oscillator.xpp")) as f:
        src.write_text(f.read())
    out = convert_xpp_file(str(src))
    with open(out) as f, open(os.path.join(GOLDEN_XPP, "van_der_pol_oscillator_reference.cl")) as g:
        assert f.read() == g.read()


def test_xpp_golden_program_compiles(tmp_path):
    from clode_amd.cpp import clode_cpp_wrapper as w

    path = os.path.join(GOLDEN_XPP, "van_der_pol_oscillator_reference.cl")
    aux = ["x2", "x3", "y4", "x5", "negx", "negy"]
    for prec in (False, True):
        s = w.FeatureSimulatorBase(w.ProblemInfo(path, ["x", "y"], ["mu"], aux, 0), "dopri5", "thresh2",
                                   w.ObserverParams(), prec, w.OpenCLResource(), "")
        s.compile()


NOISY_XPP = "init x=1\npar a=1\nwiener w1 w2\naux q = x*a\nx' = -a*x + w1 + w2\n"


def test_xpp_simulator_refuses_short_outputs(tmp_path):
    """An XPP file fixes how many aux_[] and w_[] entries getRHS touches; the
    Simulator refuses fewer names (checked before any device work)."""
    f = tmp_path / "m.xpp"
    f.write_text(NOISY_XPP)
    with pytest.raises(ValueError, match="auxiliary outputs"):
        clode.Simulator(src_file=str(f), variables={"x": 1.0}, parameters={"a": 1.0}, aux=[])
    with pytest.raises(ValueError, match="noise terms"):
        clode.Simulator(src_file=str(f), variables={"x": 1.0}, parameters={"a": 1.0}, aux=["q"], num_noise=1)


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


def _every_builtin(t: float, x: List[float], p: List[float], dx: List[float], aux: List[float],
                  w: List[float]) -> None:
    # the cases of test/test_opencl_builtins.py:62-137 (entries 26, 38, 39
    # and 40 stay 0 there: fma, logb, mad and nan are not offered)
    p0: int = int(p[0])
    p1: int = int(p[1])
    x0: float = x[0] * t
    x1: float = x[1] * t
    aux[0] = acos(x0)
    aux[1] = acosh(x0 + 1)
    aux[2] = acospi(x0)
    aux[3] = asin(x0)
    aux[4] = asinh(x0 + 1)
    aux[5] = asinpi(x0)
    aux[6] = atan(x0)
    aux[7] = atan2(x0, x1)
    aux[8] = atan2pi(x0, x1)
    aux[9] = atanh(x0 / 100.0)
    aux[10] = atanpi(x0)
    aux[11] = cbrt(x0)
    aux[12] = ceil(x0)
    aux[13] = copysign(x0, x1)
    aux[14] = cos(x0)
    aux[15] = cosh(x0)
    aux[16] = cospi(x0)
    aux[17] = erf(x0)
    aux[18] = erfc(x0)
    aux[19] = exp(x0)
    aux[20] = exp2(x0)
    aux[21] = exp10(x0)
    aux[22] = expm1(x0)
    aux[23] = fabs(x0)
    aux[24] = fdim(x0, x1)
    aux[25] = floor(x1)
    aux[26] = 0.0
    aux[27] = fmod(x0, x1 + 1)
    aux[28] = heaviside(t - 0.5)
    aux[29] = gamma(x0 + 1)
    aux[30] = hypot(x0, x1)
    aux[31] = float(ilogb(x0 + 1))
    aux[32] = ldexp(x0, int(x1))
    aux[33] = lgamma(x0 + 1)
    aux[34] = log(x0 + 1)
    aux[35] = log1p(x0 + 1)
    aux[36] = log2(x0 + 1)
    aux[37] = log10(x0 + 1)
    aux[38] = 0.0
    aux[39] = 0.0
    aux[40] = 0.0
    aux[41] = nextafter(x0, x1)
    aux[42] = pow(x0, x1)
    aux[43] = pown(x0, p0)
    aux[44] = powr(x0 + 0.5, x1 + 0.2)
    aux[45] = remainder(x0, x1 + 1)
    aux[46] = rint(x0)
    aux[47] = rootn(x0, p1)
    aux[48] = rsqrt(x0 + 1)
    aux[49] = sin(x0)
    aux[50] = sinh(x0)
    aux[51] = sinpi(x0)
    aux[52] = sqrt(x0)
    aux[53] = tan(x0)
    aux[54] = tanh(x0)
    aux[55] = tanpi(x0)
    aux[56] = trunc(x0)
    dx[0] = 0.0
    dx[1] = 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("single", [True, False])
def test_reference_builtins_case(gpu_available, single):
    """test/test_opencl_builtins.py:61-173: Euler steps of 0.5 over (0, 1),
    three stored samples, every aux value against the Python definition at
    the same time point (np.isclose with atol 1e-7 and its default rtol 1e-5)."""
    aux = [f"aux{k}" for k in range(57)]
    variables = {"x": 0.3, "y": 1.5, "z": 2.5}
    parameters = {"p0": 2.8, "p1": 4.55}
    sim = clode.TrajectorySimulator(variables=variables, parameters=parameters, rhs_equation=_every_builtin,
                                    dt=0.5, aux=aux, t_span=(0, 1), stepper=clode.Stepper.euler, max_store=3,
                                    max_steps=3, single_precision=single)
    traj = sim.trajectory()
    for k, t in enumerate([0.0, 0.5, 1.0]):
        expect = [0.0] * len(aux)
        _every_builtin(t, list(variables.values()), list(parameters.values()), [0.0] * 3, expect, [])
        for i, name in enumerate(aux):
            assert np.isclose(traj.aux[name][k], expect[i], atol=1e-7), (name, t, traj.aux[name][k], expect[i])


@pytest.mark.gpu
def test_xpp_simulator_default_outputs(gpu_available, tmp_path):
    f = tmp_path / "m.xpp"
    f.write_text(NOISY_XPP)
    sim = clode.Simulator(src_file=str(f), variables={"x": 1.0}, parameters={"a": 1.0},
                          stepper=clode.Stepper.stochastic_euler, dt=0.01, t_span=(0.0, 1.0))
    assert sim.aux_names == ["q"] and sim.num_noise == 2
    sim.seed_rng(3)
    assert np.isfinite(sim.transient(fetch_results=True)).all()


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
