"""The member-major feature matrix (clode_get_f_members, what
FeatureSimulator.get_observer_results returns): transposed on the device, it
must hold exactly the values of the feature-major getF vector, element for
element, for every precision (single widens float features to double), for
ensembles that are not a multiple of the 64-member tile and feature counts
beyond one 32-feature tile; and each result owns its array."""
import os

import numpy as np
import pytest

import clode_amd as clode
from helpers import MODELS, Pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["double", "single", "fp32c"])
@pytest.mark.parametrize("n,stored", [(1000, 0), (77, 9), (64, 3)])
def test_f_members_is_the_transpose(gpu_available, precision, n, stored):
    p = Pair("lactotroph", "dopri5", "thresh2", precision, sp=(0.1, 1.0, 1e-8, 1e-6, 10 ** 7, 10, 1),
             obs=dict(max_event_timestamps=stored), oracle=False)
    g = np.linspace(0.0, 10.0, n)
    p.set_problem_data(np.tile([[-60.0], [0.0], [0.0], [0.1]], (1, n)),
                       np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), np.linspace(2, 0, n), g]))
    p.set_tspan(0.0, 300.0)
    p.gpu.features()
    nf = p.gpu.get_n_features()
    assert nf == 46 + 2 * stored
    flat = np.array(p.gpu.get_f())
    members = np.asarray(p.gpu.get_f_members())
    assert members.shape == (n, nf) and members.flags.c_contiguous
    assert np.array_equal(members, flat.reshape((n, nf), order="F"), equal_nan=True)
    assert np.isfinite(members[:, nf - 5:]).all()  # event and step counts, dt statistics


def test_observer_results_are_independent(gpu_available):
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "van_der_pol.cl"), variables={"x": 2.0, "y": 0.0},
                                 parameters={"mu": 1.0}, observer=clode.Observer.threshold_2,
                                 stepper=clode.Stepper.rk4, dt=0.01, t_span=(0.0, 20.0))
    sim.set_ensemble(parameters={"mu": np.linspace(0.5, 3.0, 300)})
    a = sim.features().to_ndarray()
    b = sim.get_observer_results().to_ndarray()
    assert np.array_equal(a, b)
    a[:] = -1.0
    c = sim.get_observer_results().to_ndarray()
    assert np.array_equal(b, c) and not np.array_equal(a, c)
