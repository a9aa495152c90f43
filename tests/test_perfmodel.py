"""The roofline's operation table, clode_amd/perfmodel.py, is generated
mechanically: oracle/gen_perfmodel.py runs each benchmark workload's call
sequence on a member sample in the operation-counting build of the oracle
(oracle/flop_count.hpp).  These tests check that the committed table is
exactly what the generator produces from the current sources, and pin the
counting rule on steps whose operations can be counted by hand."""
import numpy as np
import pytest

import gen_perfmodel
from clode_amd import perfmodel
from oracle import OracleEnsemble


@pytest.mark.parametrize("name", list(gen_perfmodel.WORKLOADS))
def test_table_is_the_generated_count(name):
    got = gen_perfmodel.count_workload(name)
    table = perfmodel.TABLE[name]["phases"]
    assert set(got) == set(table), name
    for phase, counts in got.items():
        ref = dict(table[phase])
        per = ref.pop("per_step")
        assert counts == ref, (name, phase)
        assert gen_perfmodel.per_step(counts) == per, (name, phase)


def count_one(model, dims, stepper, sp, tspan, x0, pars):
    o = OracleEnsemble(model, *dims, stepper=stepper, precision="count", kind="base")
    o.set_solver_params(*sp)
    o.set_problem_data(x0, pars)
    o.set_tspan(tspan)
    o.transient()
    return o.counts()["transient"]


def test_counting_rule_euler_decay():
    """Euler on x' = -lambda x: x += dt k1 (mul, add), t += dt (add), and the
    wrapper's right-hand side -lambda x (mul; the negation is free): 4 FLOPs
    per step, one trial per step, nothing else."""
    c = count_one("decay", (1, 1, 0, 0), "euler", (1 / 64, 1 / 64, 1e-6, 1e-3, 10 ** 6, 10, 1), (0.0, 1.0),
                  np.ones((1, 3)), np.full((1, 3), 2.0))
    steps = c["trials"]["accepted"]
    assert steps == 3 * 65 and c["trials"]["rejected"] == 0  # t <= 1 in exact steps of 1/64
    acc = c["accepted"]
    assert (acc["add"], acc["mul"], acc["div"], acc["fma"]) == (2 * steps, 2 * steps, 0, 0)


def test_counting_rule_rk4_harmonic():
    """rk4 on x' = w y, y' = -w x (2 FLOPs per right-hand side, 4 per step):
    stage times t + h/2 and t + h (2 adds; h/2 is loop-invariant for a fixed
    step), three stages of x + c k (2 x 2 x 3), the update
    x += h (k1 + 2 k2 + 2 k3 + k4) / 6 (8 per variable): 38 FLOPs per step,
    of them 2 divisions."""
    c = count_one("harmonic", (2, 1, 0, 0), "rk4", (1 / 64, 1.0, 1e-6, 1e-3, 10 ** 6, 10, 1), (0.0, 1.0),
                  np.vstack([np.ones(2), np.zeros(2)]), np.full((1, 2), 3.0))
    steps = c["trials"]["accepted"]
    acc = c["accepted"]
    assert acc["add"] + acc["mul"] + acc["div"] + 2 * acc["fma"] == 38 * steps
    assert acc["div"] == 2 * steps


def test_launch_flops_prices_rejected_trials():
    p = perfmodel.phase_counts("lactotroph_256x256_dopri5_thresh2", "features")
    base = perfmodel.launch_flops("lactotroph_256x256_dopri5_thresh2", "features", 1000, 1000)
    assert base == pytest.approx(1000 * p["step"])
    more = perfmodel.launch_flops("lactotroph_256x256_dopri5_thresh2", "features", 1000, 1100)
    assert more - base == pytest.approx(100 * p["rejected_trial"])
    # the headline's step: a dopri5 trial with 6 right-hand sides of 3 exp each, one pow
    assert p["transcendentals"]["exp"] == pytest.approx(18, rel=0.02)
    assert p["transcendentals"]["pow"] == pytest.approx(1.0, rel=0.02)
