"""Feature simulator (counterpart of clode/features.py).

``FeatureSimulator`` runs the observer kernels (warm-up pass for two-pass
observers, then the features pass) and returns an ``ObserverOutput`` whose
structured array is indexed by feature name, as in the reference
(clode/features.py:36-143, 147-534).
"""
from __future__ import annotations

from enum import Enum
from typing import Any, Callable, Dict, List, Optional, Tuple, Union

import numpy as np
from numpy.lib import recfunctions as rfn

from .cpp.clode_cpp_wrapper import FeatureSimulatorBase, ObserverParams, SolverParams
from .runtime import CLDeviceType, CLVendor, _clode_root_dir
from .solver import Simulator, Stepper


class Observer(Enum):
    """Observers (names of clode/cpp/observers.cl:122-143)."""

    basic = "basic"
    basic_all_variables = "basicall"
    local_max = "localmax"
    neighbourhood_1 = "nhood1"
    neighbourhood_2 = "nhood2"
    threshold_2 = "thresh2"


class ObserverOutput:
    """Features of an ensemble, indexable by feature name."""

    def __init__(self, observer_params: ObserverParams, feature_array: np.ndarray, num_features: int,
                 variables: List[str], observer_type: Observer, feature_names: List[str],
                 ensemble_shape: Tuple) -> None:
        self._op = observer_params
        self._num_features = num_features
        self._vars = variables
        self._observer_type = observer_type
        self._feature_names = feature_names
        self._ensemble_shape = ensemble_shape
        dtype = np.dtype({"names": feature_names, "formats": [np.float64] * len(feature_names)})
        self.F = rfn.unstructured_to_structured(np.ascontiguousarray(feature_array), dtype=dtype)

    def __repr__(self) -> str:
        n = len(self.F[self._feature_names[0]])
        return (f"ObserverOutput( ensemble size: {n}, number of features: {len(self._feature_names)}, "
                f"feature_names: {self._feature_names})")

    def to_ndarray(self, **kwargs) -> np.ndarray:
        return rfn.structured_to_unstructured(self.F, **kwargs)

    def get_feature_names(self) -> List[str]:
        return self._feature_names

    def _get_var(self, name: str):
        try:
            result = self.F[name].squeeze().reshape(self._ensemble_shape)
        except ValueError:
            raise NotImplementedError(f"{self._observer_type} does not track {name}!")
        return result.flat[0] if result.size == 1 else result

    def get_var_max(self, var: str):
        return self._get_var(f"max {var}")

    def get_var_min(self, var: str):
        return self._get_var(f"min {var}")

    def get_var_mean(self, var: str):
        return self._get_var(f"mean {var}")

    def get_var_max_slope(self, var: str):
        return self.get_var_max(f"d{var}/dt")

    def get_var_min_slope(self, var: str):
        return self.get_var_min(f"d{var}/dt")

    def get_var_count(self, var: str):
        return self._get_var(f"{var} count")

    def get_event_data(self, name: str, type: Optional[str] = "time") -> np.ndarray:
        """Stored event data "<name> event <type> <k>", up to the first all-zero entry."""
        names = [f for f in self._feature_names if name in f and "event" in f and "count" not in f]
        if not names:
            raise NotImplementedError(f"{self._observer_type} does not track {name} event {type}s!")
        data = []
        for k in range(self._op.max_event_timestamps):
            point = self._get_var(f"{name} event {type} {k}")
            if np.all(np.asarray(point) == 0):
                break
            data.append(point)
        return np.stack(data, axis=-1).squeeze() if data else np.array([])

    def get_timestamps(self, var: str = "event") -> np.ndarray:
        first = f"{var} event time 0"
        if first not in self._feature_names:
            raise NotImplementedError(f"{self._observer_type} does not track {var} event times!")
        data = []
        for k in range(self._op.max_event_timestamps):
            point = np.asarray(self._get_var(f"{var} event time {k}"))
            if np.all(point == 0):
                break
            data.append(point[np.newaxis] if point.ndim == 0 else point)
        return np.stack(data, axis=1).squeeze() if data else np.array([])


class FeatureSimulator(Simulator):
    """Simulator that computes trajectory features on the fly (observers)."""

    _integrator: FeatureSimulatorBase

    def __init__(
        self,
        variables: Dict[str, float],
        parameters: Dict[str, float],
        aux: Optional[List[str]] = None,
        num_noise: int = 0,
        src_file: Optional[str] = None,
        rhs_equation: Optional[Callable] = None,
        supplementary_equations: Optional[List[Callable[[Any], Any]]] = None,
        stepper: Stepper = Stepper.rk4,
        dt: float = 0.1,
        dtmax: float = 1.0,
        abstol: float = 1e-6,
        reltol: float = 1e-3,
        max_steps: int = 10000000,
        max_store: int = 10000000,
        nout: int = 1,
        solver_parameters: Optional[SolverParams] = None,
        t_span: Tuple[float, float] = (0.0, 1000.0),
        single_precision: Union[bool, str] = True,
        device_type: Optional[CLDeviceType] = None,
        vendor: Optional[CLVendor] = None,
        platform_id: Optional[int] = None,
        device_id: Optional[int] = None,
        device_ids: Optional[List[int]] = None,
        observer: Observer = Observer.basic_all_variables,
        event_var: str = "",
        feature_var: str = "",
        observer_max_event_count: int = 100,
        observer_max_event_timestamps: int = 0,
        observer_min_x_amp: float = 0.0,
        observer_min_imi: float = 0.0,
        observer_neighbourhood_radius: float = 0.05,
        observer_x_up_thresh: float = 0.3,
        observer_x_down_thresh: float = 0.2,
        observer_dx_up_thresh: float = 0,
        observer_dx_down_thresh: float = 0,
        observer_eps_dx: float = 0.0,
        observer_parameters: Optional[ObserverParams] = None,
    ) -> None:
        self._observer_type = observer
        names = list(variables.keys())
        e_ix = names.index(event_var) if event_var != "" else 0
        f_ix = names.index(feature_var) if feature_var != "" else 0
        if observer_parameters is not None:
            self._op = observer_parameters
        else:
            self._op = ObserverParams(e_ix, f_ix, observer_max_event_count, observer_max_event_timestamps,
                                      observer_min_x_amp, observer_min_imi, observer_neighbourhood_radius,
                                      observer_x_up_thresh, observer_x_down_thresh, observer_dx_up_thresh,
                                      observer_dx_down_thresh, observer_eps_dx)
        self._num_features: Optional[int] = None
        super().__init__(variables=variables, parameters=parameters, aux=aux, num_noise=num_noise,
                         src_file=src_file, rhs_equation=rhs_equation,
                         supplementary_equations=supplementary_equations, stepper=stepper, dt=dt, dtmax=dtmax,
                         abstol=abstol, reltol=reltol, max_steps=max_steps, max_store=max_store, nout=nout,
                         solver_parameters=solver_parameters, t_span=t_span, single_precision=single_precision,
                         device_type=device_type, vendor=vendor, platform_id=platform_id, device_id=device_id,
                         device_ids=device_ids)

    def _create_integrator(self) -> None:
        self._integrator = FeatureSimulatorBase(self._pi, self._stepper.value, self._observer_type.value, self._op,
                                                self._single_precision, self._runtime, _clode_root_dir)

    def set_observer(self, observer_type: Observer) -> None:
        if observer_type != self._observer_type:
            self._integrator.set_observer(observer_type.value)
            self._observer_type = observer_type

    def set_observer_parameters(self, op: Optional[ObserverParams] = None, event_var: Optional[str] = None,
                                feature_var: Optional[str] = None, max_event_count: Optional[int] = None,
                                max_event_timestamps: Optional[int] = None, min_amp: Optional[float] = None,
                                min_imi: Optional[float] = None, nhood_radius: Optional[float] = None,
                                x_up_threshold: Optional[float] = None, x_down_threshold: Optional[float] = None,
                                dx_up_threshold: Optional[float] = None, dx_down_threshold: Optional[float] = None,
                                eps_dx: Optional[float] = None) -> None:
        if op is not None:
            self._op = op
        else:
            if event_var is not None:
                self._op.e_var_ix = self.variable_names.index(event_var)
            if feature_var is not None:
                self._op.f_var_ix = self.variable_names.index(feature_var)
            for name, val in (("max_event_count", max_event_count), ("max_event_timestamps", max_event_timestamps),
                              ("min_amp", min_amp), ("min_imi", min_imi), ("nhood_radius", nhood_radius),
                              ("x_up_threshold", x_up_threshold), ("x_down_threshold", x_down_threshold),
                              ("dx_up_threshold", dx_up_threshold), ("dx_down_threshold", dx_down_threshold),
                              ("eps_dx", eps_dx)):
                if val is not None:
                    setattr(self._op, name, val)
        self._integrator.set_observer_params(self._op)

    def get_observer_parameters(self) -> ObserverParams:
        return self._integrator.get_observer_params()

    def get_feature_names(self) -> List[str]:
        return self._integrator.get_feature_names()

    def is_observer_initialized(self) -> bool:
        return self._integrator.is_observer_initialized()

    def initialize_observer(self) -> None:
        """Run the observer's initialisation (warm-up pass for two-pass observers)."""
        self._integrator.initialize_observer()

    def set_replay_budget(self, nbytes: int) -> None:
        """Spend up to ``nbytes`` of device memory on the features replay:
        for a two-pass observer (threshold_2, neighbourhood_2) on a system
        without noise, the observer's warm-up pass records every member's
        accepted steps and the features pass that follows it reads them back
        instead of integrating them a second time (bit-identical results;
        8 (2 + 2 nVar + nAux) bytes per member-step, 8 nVar more in fp32c).
        0 turns it off and frees the buffer."""
        self._integrator.set_replay_budget(int(nbytes))

    def replay_info(self) -> Tuple[int, bool]:
        """(records per member of the last warm-up recording, whether the last
        features pass replayed it)."""
        cap, replayed = self._integrator.get_replay_info()
        return int(cap), bool(replayed)

    def features(self, t_span: Optional[Tuple[float, float]] = None, initialize_observer: Optional[bool] = None,
                 update_x0: bool = True, fetch_results: bool = True) -> Optional[ObserverOutput]:
        """Integrate with feature detection; optionally return the features."""
        if t_span is not None:
            self.set_tspan(t_span=t_span)
        if initialize_observer is not None:
            self._integrator.features(initialize_observer)
        else:
            self._integrator.features()
        self._device_final_state = self._device_dt = self._device_tf = None
        if update_x0:
            self._integrator.shift_x0()
            self._device_initial_state = None
        if fetch_results:
            return self.get_observer_results()
        return None

    def transient_features(self, t_transient: Tuple[float, float], t_features: Tuple[float, float],
                           update_x0: bool = True, fetch_results: bool = True) -> Optional[ObserverOutput]:
        """transient(t_transient), then features(t_features, initialize_observer=True)
        in one kernel launch: every member runs its transient, warm-up and
        features passes back to back (bit-identical to the separate calls).
        Global spans only."""
        self._integrator.transient_features(list(t_transient), list(t_features))
        self._t_span = tuple(t_features)
        self._device_final_state = self._device_dt = self._device_tf = None
        self._device_initial_state = None
        if update_x0:
            self._integrator.shift_x0()
        if fetch_results:
            return self.get_observer_results()
        return None

    def get_observer_results(self) -> ObserverOutput:
        """The features of the last launch, read from the device on every call
        (each result owns its array): (members, features) in C order,
        transposed on the device before the copy."""
        self._num_features = self._integrator.get_n_features()
        F = np.asarray(self._integrator.get_f_members())
        return ObserverOutput(self._op, F, self._num_features, self.variable_names,
                              self._observer_type, self._integrator.get_feature_names(), self._ensemble_shape)
