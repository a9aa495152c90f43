"""Trajectory simulator (counterpart of clode/trajectory.py).

Stores the initial point and every nout-th step, up to max_store samples per
member, and returns one ``TrajectoryOutput`` per member with the stored part
only (clode/trajectory.py:18-266).
"""
from __future__ import annotations

from typing import Any, Callable, Dict, List, Optional, Tuple, Union

import numpy as np
from numpy.lib import recfunctions as rfn

from .cpp.clode_cpp_wrapper import SolverParams, TrajectorySimulatorBase
from .runtime import CLDeviceType, CLVendor, _clode_root_dir
from .solver import Simulator, Stepper


class TrajectoryOutput:
    def __init__(self, t: np.ndarray, x: np.ndarray, dx: np.ndarray, aux: np.ndarray, variable_names: List[str],
                 aux_names: List[str]) -> None:
        self.t = t
        xdt = np.dtype({"names": variable_names, "formats": [np.float64] * len(variable_names)})
        self.x = rfn.unstructured_to_structured(np.ascontiguousarray(x), dtype=xdt)
        self.dx = rfn.unstructured_to_structured(np.ascontiguousarray(dx), dtype=xdt)
        if len(aux_names) > 0:
            adt = np.dtype({"names": aux_names, "formats": [np.float64] * len(aux_names)})
            self.aux = rfn.unstructured_to_structured(np.ascontiguousarray(aux), dtype=adt)
        self._variable_names = variable_names
        self._aux_names = aux_names

    def __repr__(self) -> str:
        return (f"TrajectoryOutput( length: {len(self.t)}, variable names: {self._variable_names}, "
                f"aux variable names: {self._aux_names} )")

    def to_ndarray(self, slot: str, **kwargs) -> np.ndarray:
        return rfn.structured_to_unstructured(getattr(self, slot), **kwargs)


class TrajectorySimulator(Simulator):
    """Simulator that stores trajectories."""

    _integrator: TrajectorySimulatorBase

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
        reltol: float = 1e-4,
        max_steps: int = 1000000,
        max_store: int = 1000000,
        nout: int = 1,
        solver_parameters: Optional[SolverParams] = None,
        t_span: Tuple[float, float] = (0.0, 1000.0),
        single_precision: Union[bool, str] = True,
        device_type: Optional[CLDeviceType] = None,
        vendor: Optional[CLVendor] = None,
        platform_id: Optional[int] = None,
        device_id: Optional[int] = None,
        device_ids: Optional[List[int]] = None,
    ) -> None:
        super().__init__(variables=variables, parameters=parameters, aux=aux, num_noise=num_noise,
                         src_file=src_file, rhs_equation=rhs_equation,
                         supplementary_equations=supplementary_equations, stepper=stepper, dt=dt, dtmax=dtmax,
                         abstol=abstol, reltol=reltol, max_steps=max_steps, max_store=max_store, nout=nout,
                         solver_parameters=solver_parameters, t_span=t_span, single_precision=single_precision,
                         device_type=device_type, vendor=vendor, platform_id=platform_id, device_id=device_id,
                         device_ids=device_ids)
        self._device_t = self._device_x = self._device_dx = self._device_aux = None

    def _create_integrator(self) -> None:
        self._integrator = TrajectorySimulatorBase(self._pi, self._stepper.value, self._single_precision,
                                                   self._runtime, _clode_root_dir)

    def trajectory(self, t_span: Optional[Tuple[float, float]] = None, update_x0: bool = True,
                   fetch_results: bool = True):
        """Integrate and store trajectories; optionally return them."""
        if t_span is not None:
            self.set_tspan(t_span=t_span)
        self._integrator.trajectory()
        self._device_t = self._device_x = self._device_dx = self._device_aux = None
        self._device_final_state = self._device_dt = self._device_tf = None
        if update_x0:
            self._integrator.shift_x0()
            self._device_initial_state = None
        if fetch_results:
            return self.get_trajectory()
        return None

    def get_trajectory(self):
        n_stored = np.array(self._integrator.get_n_stored())
        store = self._sp.max_store
        n, nv, na = self._ensemble_size, self.num_variables, len(self.aux_names)
        t = np.array(self._integrator.get_t(), dtype=np.float64)[: n * store].reshape((n, store), order="F")
        x = np.array(self._integrator.get_x(), dtype=np.float64)[: n * nv * store].reshape((n, nv, store),
                                                                                          order="F")
        dx = np.array(self._integrator.get_dx(), dtype=np.float64)[: n * nv * store].reshape((n, nv, store),
                                                                                            order="F")
        aux = np.array(self._integrator.get_aux(), dtype=np.float64)[: n * na * store].reshape((n, na, store),
                                                                                              order="F")
        self._device_t, self._device_x, self._device_dx, self._device_aux = t, x, dx, aux
        results = []
        for i in range(n):
            k = min(int(n_stored[i]) + 1, store)
            results.append(TrajectoryOutput(t=t[i, :k].T, x=x[i, :, :k].T, dx=dx[i, :, :k].T, aux=aux[i, :, :k].T,
                                            variable_names=self.variable_names, aux_names=self.aux_names))
        return results[0] if n == 1 else results
