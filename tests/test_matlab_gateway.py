"""The MATLAB gateways (clode_amd/matlab/clode_mex.cpp) driven in-process
through a stand-in MEX runtime (MATLAB is not installed): the reference's
action strings (matlab/clODEmex.cpp:33-91, clODEfeaturesmex.cpp:66-101,
clODEtrajectorymex.cpp), the handle table and MEX lock count, the 1-based
observer indices of the MATLAB structs, error reporting, and - on the GPU -
complete sessions whose results equal the CPU oracle's bit for bit."""
import os

import numpy as np
import pytest

from mex_fake import Gateway, MexError, problem_struct, solver_struct

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(REPO, "clode_amd", "models")
VDP = problem_struct(os.path.join(MODELS, "van_der_pol.cl"), ["x", "y"], ["mu"], [], 0)
LAC = problem_struct(os.path.join(MODELS, "lactotroph.cl"), ["v", "n", "f", "c"], ["gcal", "gsk", "gbk", "gk"],
                     ["ica"], 0)


def obs_struct(e=1, f=1, **kw):
    s = {"eVarIx": e, "fVarIx": f, "maxEventCount": 100, "minXamp": 0.0, "minIMI": 0.0, "nHoodRadius": 0.05,
         "xUpThresh": 0.3, "xDownThresh": 0.2, "dxUpThresh": 0.0, "dxDownThresh": 0.0, "eps_dx": 0.0}
    s.update(kw)
    return s


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_protocol_handles_and_lock(kind):
    g = Gateway(kind)
    args = [VDP, "rk4", 0, 0, 0] + (["thresh2"] if kind == 1 else [])
    locks = g.lock_count()
    h1, h2 = g("new", *args), g("new", *args)
    assert h1[0] == h1[0] // 1 and h2[0] == h1[0] + 1
    assert g.lock_count() == locks + 2
    assert "rk4" in g("getsteppernames", h1)
    assert "#define CLODE_STEPPER 3" in g("getprogramstring", h1)[0]
    with pytest.raises(MexError, match="Unrecognized action"):
        g("nosuchaction", h1)
    with pytest.raises(MexError, match="No instance"):
        g("transient", 999)
    with pytest.raises(MexError, match="handle"):
        g("transient")
    g("setstepper", h1, "dopri5")
    assert "#define CLODE_STEPPER 5" in g("getprogramstring", h1)[0]
    assert g("delete", h1)[0] == 0
    assert g("delete", h2)[0] == 1  # table empty
    assert g.lock_count() == locks


def test_features_gateway_observer_struct_is_one_based():
    g = Gateway(1)
    h = g("new", VDP, "rk4", 0, 0, 0, "localmax")
    assert g("getnfeatures", h)[0] == 6 + 5 * 2 + 2
    with pytest.raises(MexError, match="1-based"):
        g("setobserverpars", h, obs_struct(e=0))
    with pytest.raises(MexError):
        g("setobserverpars", h, obs_struct(e=3))  # variable 3 of 2
    g("setobserverpars", h, obs_struct(e=2, f=2, maxEventTimestamps=1))
    assert "localmax event time 0" in g("getfeaturenames", h)
    g("setobserver", h, "nhood2")
    assert g("getfeaturenames", h)[0] == "max period"
    assert set(g("getobservernames", h)) == {"basic", "basicall", "localmax", "nhood1", "nhood2", "thresh2"}
    with pytest.raises(MexError, match="too few"):
        g("setobserverpars", h)
    g("delete", h)


def test_missing_new_arguments():
    with pytest.raises(MexError):
        Gateway(1)("new", VDP, "rk4", 0, 0, 0)  # no observer
    with pytest.raises(MexError, match="clRHSfilename"):
        Gateway(0)("new", {"nVar": 2}, "rk4", 0, 0, 0)


@pytest.mark.gpu
def test_base_gateway_session_matches_oracle(gpu_available):
    from helpers import Pair

    n = 50
    mu = np.linspace(0.1, 5.0, n)
    x0 = np.vstack([2.0 * np.ones(n), np.zeros(n)])
    sp = (0.01, 1.0, 1e-6, 1e-3, 10 ** 6, 10 ** 6, 1)
    g = Gateway(0)
    h = g("new", VDP, "rk4", 0, 0, 0)
    g("setsolverpars", h, solver_struct(*sp))
    g("setproblemdata", h, x0.ravel(), mu)
    g("settspan", h, [0.0, 30.0])
    g("buildcl", h)
    g("transient", h)
    g("shiftx0", h)
    g("transient", h)
    p = Pair("van_der_pol", "rk4", "", "double", kind="base", sp=sp)
    p.set_problem_data(x0, mu[None, :])
    p.set_tspan(0.0, 30.0)
    p.orc.transient()
    p.orc.shift_x0()
    p.orc.transient()
    assert np.array_equal(g("getxf", h), p.orc.get("xf"))
    assert np.array_equal(g("gettspan", h), [0.0, 30.0])
    g("delete", h)


@pytest.mark.gpu
def test_features_gateway_session_matches_oracle(gpu_available):
    from helpers import Pair

    n = 40
    g_k = np.linspace(0.0, 10.0, n)
    x0 = np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n))
    pars = np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), np.linspace(2, 0, n), g_k])
    sp = (0.1, 1.0, 1e-8, 1e-6, 10 ** 7, 10 ** 6, 1)
    gw = Gateway(1)
    h = gw("new", LAC, "dopri5", 0, 0, 0, "thresh2")
    gw("setobserverpars", h, obs_struct())
    gw("initialize", h, [0.0, 1000.0], x0.ravel(), pars.ravel(), solver_struct(*sp))
    gw("features", h, 1)
    gw("features", h)  # continues the observer
    p = Pair("lactotroph", "dopri5", "thresh2", "double", sp=sp)
    p.set_problem_data(x0, pars)
    p.set_tspan(0.0, 1000.0)
    p.orc.features(1)
    p.orc.features(-1)
    assert np.array_equal(gw("getf", h), p.orc.get("F"))
    gw("delete", h)


@pytest.mark.gpu
def test_trajectory_gateway_session_matches_oracle(gpu_available):
    from helpers import Pair

    n = 20
    mu = np.linspace(0.5, 3.0, n)
    x0 = np.vstack([2.0 * np.ones(n), np.zeros(n)])
    sp = (0.05, 1.0, 1e-8, 1e-6, 10 ** 6, 200, 2)
    g = Gateway(2)
    h = g("new", VDP, "dopri5", 0, 0, 0)
    g("setsolverpars", h, solver_struct(*sp))
    g("setproblemdata", h, x0.ravel(), mu)
    g("settspan", h, [0.0, 20.0])
    g("trajectory", h)
    p = Pair("van_der_pol", "dopri5", "", "double", kind="trajectory", sp=sp)
    p.set_problem_data(x0, mu[None, :])
    p.set_tspan(0.0, 20.0)
    p.orc.trajectory()
    ns = g("getnstored", h).astype(int)
    assert np.array_equal(ns, p.orc.get("nStored").astype(int))
    t, x = g("gett", h).reshape(200, n), g("getx", h).reshape(200, 2, n)
    rt, rx = p.orc.get("t").reshape(200, n), p.orc.get("x").reshape(200, 2, n)
    for i in range(n):
        k = min(ns[i], 200)
        assert np.array_equal(t[:k, i], rt[:k, i]) and np.array_equal(x[:k, :, i], rx[:k, :, i])
    g("delete", h)
