"""The C ABI library loads on a host without a GPU and exports every symbol
include/drift_hip.h declares.  Only host-only entry points are called."""

from __future__ import annotations

import ctypes
import re

import pytest

from astrodrift_amd import _native


def declared_functions() -> list[str]:
    text = _native.HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?[a-z_]+\**\s+\**(drift_[a-z0-9_]+)\s*\(", text, flags=re.M)


def test_header_declares_the_expected_entry_points():
    names = set(declared_functions())
    assert {
        "drift_get_abi_version", "drift_last_error", "drift_session_create",
        "drift_session_destroy", "drift_galpy_dopr54", "drift_galpy_dop853",
        "drift_galpy_dopr54_device", "drift_galpy_dop853_device", "drift_force_device",
        "drift_mw_table_check", "drift_origins_create", "drift_origins_destroy",
        "drift_origin_tracks_hermite", "drift_server_create", "drift_server_submit",
        "drift_server_query", "drift_server_wait", "drift_server_stats", "drift_server_destroy",
        "drift_ragged_device", "drift_tune_ragged", "drift_tune_device",
        "drift_launch_shape", "drift_session_last_launch", "drift_server_busy",
        "drift_galpy_dop853_device_noalloc",
    } <= names


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_abi_version_and_struct_layouts():
    lib = _native.load_library()
    assert lib.drift_get_abi_version() == _native.ABI_VERSION == 6
    assert ctypes.sizeof(_native.CPotential) == 8 + 10 * 8
    assert ctypes.sizeof(_native.CLaunchReport) == 4 * 8 + 2 * 8 + 2 * 4
    assert ctypes.sizeof(_native.COrbitStats) == 16
    assert ctypes.sizeof(_native.CLaunchOptions) == 24
    assert _native.STATS_DTYPE.itemsize == 16


def test_table_check_validates_its_arguments():
    with pytest.raises(_native.DriftNativeError):
        _native.mw_table_check(_native.PotentialSpec.kepler(1.0))
    bad = _native.PotentialSpec.milky_way(bulge_alpha=3.5)
    with pytest.raises(_native.DriftNativeError, match="alpha"):
        _native.mw_table_check(bad)


def test_null_session_is_an_error_not_a_crash():
    lib = _native.load_library()
    rc = lib.drift_galpy_dop853(None, None, None, None, 1, 2, 0.0, 0.0, None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT
    assert "session" in _native.last_error()


def test_null_handles_of_the_tuners_and_server_are_errors():
    """Argument checks that run before any device call."""
    import ctypes

    lib = _native.load_library()
    from astrodrift_amd import serving

    serving._declare(lib)
    rc = lib.drift_tune_ragged(None, None, 0, None, 0, None, 0, None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and "session" in _native.last_error()
    rc = lib.drift_tune_device(None, 1, None, None, 10, 10, -23.0, -23.0, -9999.99, None, 0,
                               None, None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and "session" in _native.last_error()
    g, b = ctypes.c_int32(3), ctypes.c_int32(4)
    rc = lib.drift_launch_shape(None, 1, None, 10, None, ctypes.byref(g), ctypes.byref(b))
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and (g.value, b.value) == (3, 4)
    rc = lib.drift_session_last_launch(None, ctypes.byref(_native.CLaunchReport()))
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and "NULL" in _native.last_error()
    v = ctypes.c_uint64(7)
    rc = lib.drift_server_stats(None, ctypes.byref(v), None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and v.value == 7
    assert "server" in _native.last_error()
    busy = ctypes.c_double(2.5)
    rc = lib.drift_server_busy(None, ctypes.byref(busy), None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and busy.value == 2.5
