"""clode_amd: MI355X-native ensemble ODE/SDE integration with on-the-fly
feature extraction.

Public interface of the reference's ``clode`` package (clode/__init__.py):
``Simulator``, ``FeatureSimulator``, ``TrajectorySimulator``, ``Stepper``,
``Observer``, ``ObserverOutput``, ``TrajectoryOutput``, the parameter structs
and the runtime/logging helpers.  The compute path is the native library
``libclode_hip.so`` (hipRTC-compiled gfx950 kernels behind the C ABI in
include/clode_hip.h); importing this package fails if the native extension
has not been built (``python -m clode_amd.build``).
"""
from __future__ import annotations

from .cpp.clode_cpp_wrapper import ObserverParams, ProblemInfo, SolverParams
from .batch import BatchSimulator
from .features import FeatureSimulator, Observer, ObserverOutput
from .function_converter import OpenCLConverter, OpenCLRhsEquation, convert_str_to_opencl
from .opencl_builtins import *  # noqa: F401,F403  (the built-ins a Python right-hand side may call)
from .runtime import (
    CLDeviceType,
    CLVendor,
    DeviceInfo,
    LogLevel,
    PlatformInfo,
    get_log_level,
    initialize_runtime,
    print_opencl,
    query_opencl,
    set_log_level,
    set_log_pattern,
)
from .solver import Simulator, Stepper
from .trajectory import TrajectoryOutput, TrajectorySimulator
from .xpp_parser import convert_xpp_file

__version__ = "0.1.0"

__all__ = [
    "OpenCLConverter",
    "OpenCLRhsEquation",
    "convert_str_to_opencl",
    "convert_xpp_file",
    "BatchSimulator",
    "CLDeviceType",
    "CLVendor",
    "DeviceInfo",
    "PlatformInfo",
    "initialize_runtime",
    "print_opencl",
    "query_opencl",
    "LogLevel",
    "get_log_level",
    "set_log_level",
    "set_log_pattern",
    "ProblemInfo",
    "SolverParams",
    "Stepper",
    "Simulator",
    "FeatureSimulator",
    "Observer",
    "ObserverParams",
    "ObserverOutput",
    "TrajectorySimulator",
    "TrajectoryOutput",
]
