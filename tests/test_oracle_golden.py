"""Pins the oracle against values the reference itself printed
(tests/golden/reference_values.json)."""
import json
import os

import numpy as np

from oracle import OracleEnsemble

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def test_lactotroph_transient_fp32_golden():
    g = GOLDEN["lactotroph_transient_fp32"]
    o = OracleEnsemble("lactotroph", 4, 4, 1, stepper="rk4", precision="single", kind="base", threads=1)
    o.set_solver_params(g["dt"], 1.0, 1e-6, 1e-3, 10000000, 10000000, 50)
    n = 4
    p = g["pars"]
    o.set_problem_data(np.zeros((4, n)), np.array([[p["gcal"]] * n, [p["gsk"]] * n, [p["gbk"]] * n, [p["gk"]] * n]))
    o.set_tspan(g["tspan"])
    o.transient()
    xf = o.get("xf").reshape(4, n)[:, 0]
    # 6 printed significant digits; the tiny BK gate f agrees to 5 (device
    # math libraries differ in the last float bits)
    np.testing.assert_allclose(xf[[0, 1, 3]], np.array(g["xf"])[[0, 1, 3]], rtol=1e-5)
    np.testing.assert_allclose(xf[2], g["xf"][2], rtol=2e-5)
    assert abs(o.get("tf")[0] - g["tf"]) < 0.1
