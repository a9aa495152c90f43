"""Ensemble simulator base class (counterpart of clode/solver.py).

``Simulator`` owns one native simulator (``SimulatorBase`` from the pybind11
module over include/clode_hip.h) and adds the reference's ensemble
handling: default values, broadcasting of scalar/array parameter and
initial-state specifications (clode/solver.py:254-462), column-major
flattening at the boundary and lazy device-to-host transfers.
"""
from __future__ import annotations

import os
from enum import Enum
from typing import Any, Callable, Dict, List, Mapping, Optional, Tuple, Union

import numpy as np

from .cpp.clode_cpp_wrapper import ProblemInfo, SimulatorBase, SolverParams
from .runtime import CLDeviceType, CLVendor, OpenCLResource, _clode_root_dir, initialize_runtime


class Stepper(Enum):
    """Time steppers (registry of clode/cpp/steppers.cl:26-43)."""

    euler = "euler"
    heun = "heun"
    rk4 = "rk4"
    bs23 = "bs23"
    dormand_prince = "dopri5"
    stochastic_euler = "seuler"
    heun_euler = "he12"  # adaptive_he12.clh (not in the reference's name map)


def _precision_arg(single_precision: Union[bool, str]) -> Union[bool, str]:
    if isinstance(single_precision, str):
        if single_precision not in ("single", "double", "fp32c"):
            raise ValueError("precision must be 'single', 'double' or 'fp32c'")
    return single_precision


def _write_rhs_source(rhs_equation, supplementary_equations) -> str:
    from .function_converter import OpenCLConverter

    converter = OpenCLConverter()
    if supplementary_equations is not None:
        for eq in supplementary_equations:
            converter.convert_to_opencl(eq)
    text = converter.convert_to_opencl(rhs_equation, mutable_args=[3, 4], function_name="getRHS")
    # the program builder reads the right-hand side from a file; name it by
    # content so that equal sources share a file (and a compiled program)
    import hashlib
    import tempfile

    digest = hashlib.sha1(text.encode()).hexdigest()[:16]
    path = os.path.join(tempfile.gettempdir(), f"clode_rhs_{digest}.cl")
    with open(path, "w") as f:
        f.write(text)
    return path


class Simulator:
    """Simulate an ensemble of instances of an ODE/SDE system, keeping only the
    final state (transient integration)."""

    _integrator: SimulatorBase
    _ensemble_size: int
    _ensemble_shape: Tuple

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
        self._xpp_outputs: Optional[Tuple[List[str], List[str]]] = None
        input_file = self._handle_clode_rhs_cl_file(src_file, rhs_equation, supplementary_equations)
        aux, num_noise = self._check_xpp_outputs(aux, num_noise)
        self._pi = ProblemInfo(input_file, list(variables.keys()), list(parameters.keys()),
                               list(aux) if aux is not None else [], num_noise)
        self._stepper = stepper
        self._single_precision = _precision_arg(single_precision)
        self._runtime: OpenCLResource = initialize_runtime(device_type, vendor, platform_id, device_id,
                                                           device_ids)
        self._device_parameters: Optional[np.ndarray] = None
        self._device_initial_state: Optional[np.ndarray] = None
        self._device_final_state: Optional[np.ndarray] = None
        self._device_dt: Optional[np.ndarray] = None
        self._device_tf: Optional[np.ndarray] = None
        self._create_integrator()
        self._build_cl_program()

        self._sp = solver_parameters if solver_parameters is not None else SolverParams(
            dt, dtmax, abstol, reltol, max_steps, max_store, nout)
        self.set_solver_parameters()
        self.set_tspan(t_span=t_span)

        self._variable_defaults = dict(variables)
        self._parameter_defaults = dict(parameters)
        self._ensemble_size = 1
        self._ensemble_shape = (1,)
        x0 = np.array(list(self._variable_defaults.values()), dtype=np.float64, ndmin=2)
        p = np.array(list(self._parameter_defaults.values()), dtype=np.float64, ndmin=2)
        self._set_problem_data(x0, p)

    # ------------------------------------------------------------- properties
    @property
    def variable_names(self) -> List[str]:
        return self._pi.vars

    @property
    def num_variables(self) -> int:
        return self._pi.num_var

    @property
    def parameter_names(self) -> List[str]:
        return self._pi.pars

    @property
    def num_parameters(self) -> int:
        return self._pi.num_par

    @property
    def aux_names(self) -> List[str]:
        return self._pi.aux

    @property
    def num_aux(self) -> int:
        return self._pi.num_aux

    @property
    def num_noise(self) -> int:
        return self._pi.num_noise

    # ------------------------------------------------------------ construction
    def _create_integrator(self) -> None:
        self._integrator = SimulatorBase(self._pi, self._stepper.value, self._single_precision, self._runtime,
                                         _clode_root_dir)

    def _build_cl_program(self) -> None:
        self._integrator.build_cl()

    def _handle_clode_rhs_cl_file(self, src_file, rhs_equation, supplementary_equations) -> str:
        if src_file is not None and rhs_equation is not None:
            raise ValueError("Cannot specify both src_file and rhs_equation")
        if src_file is not None:
            if src_file.endswith(".xpp") or src_file.endswith(".ode"):
                from .xpp_parser import convert_xpp_file, read_ode_parameters

                with open(src_file) as f:
                    parsed = read_ode_parameters(f.read())
                self._xpp_outputs = (list(parsed[1]), list(parsed[4]))
                return convert_xpp_file(src_file)
            return src_file
        if rhs_equation is not None:
            return _write_rhs_source(rhs_equation, supplementary_equations)
        raise ValueError("Must specify either src_file or rhs_equation")

    def _check_xpp_outputs(self, aux: Optional[List[str]], num_noise: int) -> Tuple[Optional[List[str]], int]:
        """An XPP file fixes how many aux_[] entries getRHS writes and how many
        w_[] entries it reads: default the aux names and noise count to the
        file's, and refuse smaller ones (the kernel would index past its
        arrays)."""
        if self._xpp_outputs is None:
            return aux, num_noise
        file_aux, file_noise = self._xpp_outputs
        if aux is None:
            aux = file_aux
        elif len(aux) < len(file_aux):
            raise ValueError(f"The XPP file defines {len(file_aux)} auxiliary outputs ({', '.join(file_aux)}); "
                             f"aux names {len(aux)}")
        if num_noise == 0:
            num_noise = len(file_noise)
        elif num_noise < len(file_noise):
            raise ValueError(f"The XPP file uses {len(file_noise)} noise terms; num_noise is {num_noise}")
        return aux, num_noise

    # ---------------------------------------------------------------- ensembles
    def set_repeat_ensemble(self, num_repeats: int) -> None:
        """Ensemble of num_repeats identical copies of the current (or default) member."""
        x0, p = self._make_problem_data(new_size=num_repeats, new_shape=(num_repeats, 1))
        self._set_problem_data(x0, p)

    def set_ensemble(
        self,
        variables: Optional[Union[np.ndarray, Mapping[str, Any]]] = None,
        parameters: Optional[Union[np.ndarray, Mapping[str, Any]]] = None,
    ) -> None:
        """Set initial states and/or parameters, possibly changing the ensemble size.

        Arrays are (ensemble_size, num_variables) / (ensemble_size, num_parameters);
        mappings give per-name scalars (broadcast) or equal-shape arrays (the
        array shape becomes the ensemble shape).  Unspecified values come from
        the current ensemble when its size is 1 or unchanged, else from the
        defaults."""
        if variables is None and parameters is None:
            raise ValueError("initial_state and parameters cannot both be None")
        var_size, var_shape, variables = self._check_spec(variables, self.variable_names, "variables")
        par_size, par_shape, parameters = self._check_spec(parameters, self.parameter_names, "parameters")
        if var_size > 1 and par_size > 1 and var_size != par_size:
            raise ValueError("Arrays specified for parameters and initial states must have the same size")
        new_size = var_size if var_size > 1 else par_size
        new_shape = var_shape if var_size > 1 else par_shape
        x0, p = self._make_problem_data(variables=variables, parameters=parameters, new_size=new_size,
                                        new_shape=new_shape)
        self._set_problem_data(x0, p)

    @staticmethod
    def _check_spec(spec, names: List[str], what: str):
        size, shape = 1, (1,)
        if spec is None:
            return size, shape, spec
        if isinstance(spec, np.ndarray):
            if spec.ndim != 2 or spec.shape[1] != len(names):
                raise ValueError(f"{'initial_state' if what == 'variables' else what} must be a matrix with "
                                 f"{len(names)} columns")
            return spec.shape[0], (spec.shape[0], 1), spec
        if isinstance(spec, Mapping):
            unknown = set(spec.keys()) - set(names)
            if unknown:
                raise ValueError(f"Unknown {what[:-1]} name(s): {unknown}")
            arrays = {k: np.array(v, dtype=np.float64) for k, v in spec.items()}
            shapes = {k: v.shape for k, v in arrays.items() if v.size > 1}
            if len(set(shapes.values())) > 1:
                raise ValueError(f"Shape of arrays for {what} don't match: {shapes}")
            if shapes:
                shape = next(iter(shapes.values()))
                size = int(np.prod(shape))
            return size, shape, arrays
        raise ValueError(f"Expected np.ndarray or Mapping for {what}, but got {type(spec)}")

    def _make_problem_data(self, variables=None, parameters=None, new_size: int = 1,
                           new_shape: Optional[Tuple] = None):
        if new_shape is None or len(new_shape) == 1:
            new_shape = (new_size, 1)
        previous = self._ensemble_size
        if previous == new_size or previous == 1:
            x0 = np.array(self.get_initial_state(), dtype=np.float64, copy=True)
            p = np.array(self._device_parameters, dtype=np.float64, copy=True)
        else:
            x0 = np.array(list(self._variable_defaults.values()), dtype=np.float64, ndmin=2)
            p = np.array(list(self._parameter_defaults.values()), dtype=np.float64, ndmin=2)
        if x0.shape[0] == 1:
            x0 = np.tile(x0, (new_size, 1))
        if p.shape[0] == 1:
            p = np.tile(p, (new_size, 1))
        if isinstance(variables, np.ndarray):
            x0 = np.array(variables, dtype=np.float64)
        elif isinstance(variables, Mapping):
            for k, v in variables.items():
                v = np.repeat(v, new_size) if v.size == 1 else v
                x0[:, self.variable_names.index(k)] = np.asarray(v).flatten()
        if isinstance(parameters, np.ndarray):
            p = np.array(parameters, dtype=np.float64)
        elif isinstance(parameters, Mapping):
            for k, v in parameters.items():
                v = np.repeat(v, new_size) if v.size == 1 else v
                p[:, self.parameter_names.index(k)] = np.asarray(v).flatten()
        self._ensemble_size = new_size
        self._ensemble_shape = new_shape
        return x0, p

    def _set_problem_data(self, initial_state: np.ndarray, parameters: np.ndarray) -> None:
        self._device_initial_state = initial_state
        self._device_parameters = parameters
        self._integrator.set_problem_data(initial_state.flatten(order="F"), parameters.flatten(order="F"))
        self._device_final_state = self._device_dt = self._device_tf = None

    def _set_parameters(self, parameters: np.ndarray) -> None:
        self._device_parameters = parameters
        self._integrator.set_pars(parameters.flatten(order="F"))

    def _set_initial_state(self, initial_state: np.ndarray) -> None:
        self._device_initial_state = initial_state
        self._integrator.set_x0(initial_state.flatten(order="F"))

    # ------------------------------------------------------------ parameters
    def set_tspan(self, t_span: Tuple[float, float]) -> None:
        self._t_span = tuple(t_span)
        self._integrator.set_tspan(list(t_span))

    def get_tspan(self) -> Tuple[float, float]:
        self._t_span = tuple(self._integrator.get_tspan())
        return self._t_span

    def shift_tspan(self) -> None:
        self._integrator.shift_tspan()
        self._t_span = tuple(self._integrator.get_tspan())

    def set_solver_parameters(self, solver_parameters: Optional[SolverParams] = None, dt=None, dtmax=None,
                              abstol=None, reltol=None, max_steps=None, max_store=None, nout=None) -> None:
        if solver_parameters is not None:
            self._sp = solver_parameters
        else:
            for name, val in (("dt", dt), ("dtmax", dtmax), ("abstol", abstol), ("reltol", reltol),
                              ("max_steps", max_steps), ("max_store", max_store), ("nout", nout)):
                if val is not None:
                    setattr(self._sp, name, val)
        self._integrator.set_solver_params(self._sp)

    def get_solver_parameters(self) -> SolverParams:
        return self._integrator.get_solver_params()

    def seed_rng(self, seed: Optional[int] = None) -> None:
        if seed is not None:
            self._integrator.seed_rng(seed)
        else:
            self._integrator.seed_rng()

    # ----------------------------------------------------------- simulation
    def transient(self, t_span: Optional[Tuple[float, float]] = None, update_x0: bool = True,
                  fetch_results: bool = False) -> Optional[np.ndarray]:
        """Integrate over the time span keeping only the final state."""
        if t_span is not None:
            self.set_tspan(t_span=t_span)
        self._integrator.transient()
        self._device_final_state = self._device_dt = self._device_tf = None
        if update_x0:
            self._integrator.shift_x0()
            self._device_initial_state = None
        if fetch_results:
            return self.get_final_state()
        return None

    # ------------------------------------------------------------- results
    def get_initial_state(self) -> np.ndarray:
        if self._device_initial_state is None:
            self._device_initial_state = np.array(self._integrator.get_x0(), dtype=np.float64).reshape(
                (self._ensemble_size, self.num_variables), order="F")
        return self._device_initial_state

    def get_final_state(self) -> np.ndarray:
        if self._device_final_state is None:
            self._device_final_state = np.array(self._integrator.get_xf(), dtype=np.float64).reshape(
                (self._ensemble_size, self.num_variables), order="F")
        return self._device_final_state

    def get_dt(self) -> np.ndarray:
        if self._device_dt is None:
            self._device_dt = np.array(self._integrator.get_dt(), dtype=np.float64).reshape(
                self._ensemble_shape, order="F")
        return self._device_dt

    def get_final_time(self) -> np.ndarray:
        if self._device_tf is None:
            self._device_tf = np.array(self._integrator.get_tf(), dtype=np.float64).reshape(
                self._ensemble_shape, order="F")
        return self._device_tf

    # ---------------------------------------------------------------- device
    def get_max_memory_alloc_size(self, deviceID: int = 0) -> int:
        return self._runtime.get_max_memory_alloc_size(deviceID)

    def get_double_support(self, deviceID: int = 0) -> bool:
        return self._runtime.get_double_support(deviceID)

    def get_device_cl_version(self, deviceID: int = 0) -> str:
        return self._runtime.get_device_cl_version(deviceID)

    def get_available_steppers(self) -> List[str]:
        return self._integrator.get_available_steppers()

    def get_program_string(self) -> str:
        return self._integrator.get_program_string()

    def print_status(self) -> None:
        self._integrator.print_status()

    def print_devices(self) -> None:
        self._runtime.print_devices()

    # --------------------------------------------------- extensions (native)
    def set_launch(self, mode: str = "grid", block_size: int = 0, slice: Optional[int] = None) -> None:
        """Work distribution: 'grid' (one member per lane) or 'persistent'
        (resident lanes claim the next member from an atomic work counter as
        soon as theirs finishes; ``slice`` loop iterations between checks).
        Results are identical in both modes."""
        self._integrator.set_launch({"grid": 0, "persistent": 1}[mode], block_size)
        if slice is not None:
            self._integrator.set_queue_slice(int(slice))

    def set_member_tspan(self, t_start, t_end=None) -> None:
        """Per-member time spans (no reference counterpart).  ``t_start`` and
        ``t_end`` broadcast to the ensemble shape; ``set_member_tspan(None)``
        returns to the global span.  The start times replace ``t_span[0]`` and
        the end times ``t_span[1]`` member by member in every launch."""
        if t_start is None:
            self._integrator.set_member_tspan([])
            return
        n = self._ensemble_size

        def flat(t):
            t = np.asarray(t, dtype=np.float64)
            if t.size == n:
                return t.flatten(order="F")
            return np.broadcast_to(t, self._ensemble_shape).flatten(order="F")

        t0, t1 = flat(t_start), flat(t_end)
        if t0.size != n or t1.size != n:
            raise ValueError("member time spans must broadcast to the ensemble shape")
        self._integrator.set_member_tspan(list(np.concatenate([t0, t1])))

    def last_kernel_ms(self) -> float:
        """Duration of the last kernel launch (HIP events), milliseconds."""
        return self._integrator.last_kernel_ms()

    def last_launch_variant(self) -> int:
        """0 when the last launch ran the default compilation of the program;
        w > 0 when it ran the variant compiled for at least w waves per SIMD
        (grid launches of ensembles larger than the default kernel keeps
        resident)."""
        return self._integrator.last_launch_variant()

    def enable_step_counter(self, on: bool = True) -> None:
        """Count accepted steps per member over all launches (member-steps)."""
        self._integrator.set_step_counter(on)

    def reset_step_counts(self) -> None:
        self._integrator.reset_step_counts()

    def get_step_counts(self) -> np.ndarray:
        return np.array(self._integrator.get_step_counts(), dtype=np.uint64).reshape(self._ensemble_shape,
                                                                                     order="F")

    def get_trial_counts(self) -> np.ndarray:
        """Step trials per member over the counted launches: the accepted
        steps plus the adaptive controller's rejected trials."""
        return np.array(self._integrator.get_trial_counts(), dtype=np.uint64).reshape(self._ensemble_shape,
                                                                                      order="F")
