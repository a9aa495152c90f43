"""Observer continuation, pinned by mathematics.

Reference: a features() call after the first one continues the observer
(clode/cpp/CLODEfeatures.cpp:215-229); at the end of every call each observer
shifts its stored times by -(tf - t0) (finalizeObserverData,
observers/observer_threshold_2.clh:458-470, observer_local_maximum.clh:
319-328, observer_neighborhood_2.clh:305-313), so that the next call, which
starts again at tspan[0] from x0 <- xf, finds them relative to its own start.

Setup that makes the answer exact: autonomous right-hand sides (harmonic
oscillator, Van der Pol), fixed-step rk4 with dt = 1/64 and spans that end
half a step past a multiple of dt.  Every step time, event time and time
difference is then an exact dyadic rational, and the states do not depend on
absolute time.  So one features() over (0, Ta + Tb) must equal, bit for bit,
features() over (0, Ta), shift_x0(), and a continued features() over
(0, Tb): the same steps, the same events, the same time differences.  Both
runs take their thresholds from one warm-up over the whole span
(initialize_observer before the first part).  Periods, up/down durations,
duty cycles, inter-maximum intervals, step sizes and time means all involve
differences of stored and current times, so an observer that dropped the
shift, shifted the wrong field or by the wrong amount, fails; the planted
mutant (the oracle built without the shift) shows that it does.
"""
import numpy as np
import pytest

from helpers import MODEL_INFO, MODELS
from oracle import OracleEnsemble

DT = 1.0 / 64
TA, TB = 17.0, 23.0
SP = (DT, 1.0, 1e-6, 1e-3, 10 ** 7, 1000, 1)

OBS = dict(e_var_ix=0, f_var_ix=0, max_event_count=12, max_event_timestamps=0, min_amp=0.0, min_imi=0.0,
           nhood_radius=0.05, x_up_threshold=0.3, x_down_threshold=0.2, dx_up_threshold=0.0,
           dx_down_threshold=0.0, eps_dx=0.0)


def problem(model, n=9):
    if model == "harmonic":
        return np.vstack([np.ones(n), np.zeros(n)]), np.linspace(1.5, 3.1, n)[None, :]
    return np.vstack([2.0 * np.ones(n), np.zeros(n)]), np.linspace(0.3, 3.0, n)[None, :]


# event limits that some members reach in the second part and others never
MAX_EVENTS = {"harmonic": 12, "van_der_pol": 5}


def obs_params(model):
    return dict(OBS, max_event_count=MAX_EVENTS[model])


def ensemble(model, observer, variant=""):
    nv, npar, na = (len(v) for v in MODEL_INFO[model][:3])
    o = OracleEnsemble(model, nv, npar, na, 0, stepper="rk4", observer=observer, threads=1, variant=variant)
    o.set_solver_params(*SP)
    o.set_observer_params(**obs_params(model))
    x0, pars = problem(model)
    o.set_problem_data(x0, pars)
    return o


def span_to(T):
    return (0.0, T - DT / 2)  # the t <= t1 loop ends exactly at T


def long_run(model, observer, variant=""):
    o = ensemble(model, observer, variant)
    o.set_tspan(span_to(TA + TB))
    o.features(1)
    return o.get("F"), o


def split_run(model, observer, variant=""):
    o = ensemble(model, observer, variant)
    o.set_tspan(span_to(TA + TB))
    o.initialize_observer()
    o.set_tspan(span_to(TA))
    o.features()  # continues the initialised observer
    first = o.get("F")
    assert np.all(o.get("tf") == TA)
    o.shift_x0()
    o.set_tspan(span_to(TB))
    o.features()
    return o.get("F"), first, o


def names(model, observer):
    from clode_amd.cpp import clode_cpp_wrapper as w

    vars_, pars, aux, noise = MODEL_INFO[model]
    pi = w.ProblemInfo(f"{MODELS}/{model}.cl", vars_, pars, aux, noise)
    sim = w.FeatureSimulatorBase(pi, "rk4", observer, w.ObserverParams(**OBS), "double", w.OpenCLResource(), "")
    return list(sim.get_feature_names())


@pytest.mark.parametrize("observer", ["thresh2", "localmax", "nhood2"])
@pytest.mark.parametrize("model", ["harmonic", "van_der_pol"])
def test_continued_features_equal_one_long_run(model, observer):
    F_long, o_long = long_run(model, observer)
    F_split, F_first, o_split = split_run(model, observer)
    n = 9
    nf = F_long.size // n
    F_long, F_split, F_first = (a.reshape(nf, n) for a in (F_long, F_split, F_first))
    same = F_long == F_split
    assert same.all(), f"{int((~same).sum())} of {F_long.size} features differ: " + ", ".join(
        sorted({names(model, observer)[k] for k in np.argwhere(~same)[:, 0]}))
    nm = names(model, observer)
    steps = F_long[nm.index("step count")]
    events = F_long[nm.index("event count")]
    first_events = F_first[nm.index("event count")]
    # every member stepped through both parts unless it reached the event limit
    assert np.all((steps == 64 * (TA + TB)) | (events == MAX_EVENTS[model]))
    # events in both parts, so for every member one period spans the cut
    assert np.all(first_events >= 1) and np.all(events > first_events)
    assert np.array_equal(o_long.get("xf"), o_split.get("xf"))
    # the split run ends at Tb in its own time, Ta earlier than the long run
    tf_long, tf_split = o_long.get("tf"), o_split.get("tf")
    assert np.array_equal(tf_long, tf_split + TA)
    assert np.all((tf_split == TB) | (events == MAX_EVENTS[model]))
    if observer != "localmax":  # the two-pass observers: terminal events in the second part
        assert np.any((events == MAX_EVENTS[model]) & (first_events < MAX_EVENTS[model]))


@pytest.mark.parametrize("observer", ["thresh2", "localmax", "nhood2"])
def test_missing_time_shift_is_detected(observer):
    """The oracle built without finalizeObserverData's shift (a planted
    mistake, CPU only) still integrates the same states but its continued
    features differ from the long run, so the test above would fail."""
    F_long, _ = long_run("harmonic", observer, variant="noshift")
    F_split, _, _ = split_run("harmonic", observer, variant="noshift")
    assert not np.array_equal(F_long, F_split)
    F_good, _ = long_run("harmonic", observer)
    assert np.array_equal(F_long, F_good)  # a single call never shifts before it reads
