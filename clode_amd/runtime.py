"""Device selection and logging (counterpart of clode/runtime.py).

The reference selects an OpenCL platform/device; here ``OpenCLResource``
selects a HIP device ordinal.  The argument checks of
``initialize_runtime`` raise the same ``ValueError``s as the reference
(clode/runtime.py:22-52).
"""
from __future__ import annotations

import os

from .cpp.clode_cpp_wrapper import (
    CLDeviceType,
    CLVendor,
    DeviceInfo,
    LogLevel,
    OpenCLResource,
    PlatformInfo,
    _print_opencl,
    get_logger,
    query_opencl,
)

#: directory of the on-disk hipRTC code-object cache (passed as clode_root)
_clode_root_dir: str = os.environ.get(
    "CLODE_CACHE_DIR", os.path.join(os.path.dirname(os.path.abspath(__file__)), "_jitcache")
)

DEFAULT_LOG_LEVEL = LogLevel.warn


def initialize_runtime(
    device_type: CLDeviceType | None,
    vendor: CLVendor | None,
    platform_id: int | None,
    device_id: int | None,
    device_ids: list[int] | None,
) -> OpenCLResource:
    """Return a device selector; mirrors the reference's validation rules."""
    if platform_id is not None:
        if device_type is not None:
            raise ValueError("Cannot specify device_type when platform_id is specified")
        if vendor is not None:
            raise ValueError("Cannot specify vendor when platform_id is specified")
        if device_id is not None and device_ids is not None:
            raise ValueError("Cannot specify both device_id and device_ids")
        if device_id is None and device_ids is None:
            raise ValueError("Must specify one of device_id and device_ids")
        if device_id is not None:
            return OpenCLResource(platform_id, device_id)
        return OpenCLResource(platform_id, device_ids)
    if device_id is not None:
        raise ValueError("Must specify platform_id when specifying device_id")
    if device_ids is not None:
        raise ValueError("Must specify platform_id when specifying device_ids")
    if device_type is None:
        device_type = CLDeviceType.DEVICE_TYPE_DEFAULT
    if vendor is None:
        vendor = CLVendor.VENDOR_ANY
    return OpenCLResource(device_type, vendor)


def get_log_level() -> LogLevel:
    return get_logger().get_log_level()


def set_log_level(level: LogLevel) -> None:
    get_logger().set_log_level(level)


def set_log_pattern(pattern: str) -> None:
    get_logger().set_log_pattern(pattern)


set_log_level(DEFAULT_LOG_LEVEL)


def print_opencl() -> None:
    old = get_log_level()
    set_log_level(LogLevel.info)
    _print_opencl()
    set_log_level(old)


__all__ = [
    "CLDeviceType",
    "CLVendor",
    "DeviceInfo",
    "PlatformInfo",
    "OpenCLResource",
    "initialize_runtime",
    "print_opencl",
    "query_opencl",
    "DEFAULT_LOG_LEVEL",
    "LogLevel",
    "set_log_level",
    "set_log_pattern",
    "get_log_level",
]
