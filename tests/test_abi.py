"""CPU checks of the C ABI library: it loads, exports every declared entry
point, and validates configurations before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import birdflow_hip as bh
from birdflow_hip import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "birdflow_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bf_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_binding():
    assert _declared_functions() == sorted(_abi.EXPORTED_FUNCTIONS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIBRARY_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name
    assert _abi.load().bf_abi_version() == 1


def test_config_struct_size_matches_header():
    # 10 uint32 + 6 uint32 boundary + 3+1+1+1+1 + 2 + 2 + 3 + 1 + 3 floats + arithmetic + 7 reserved
    assert ctypes.sizeof(_abi.Config) == 4 * (10 + 6 + 3 + 1 + 1 + 1 + 1 + 2 + 2 + 3 + 1 + 3 + 1 + 7)
    assert _abi.Config.arithmetic.offset == 4 * (10 + 6 + 3 + 1 + 1 + 1 + 1 + 2 + 2 + 3 + 1 + 3)


@pytest.mark.parametrize("field,value", [("nx", 0), ("omega_plus", 2.5), ("storage", 7), ("force_mode", 9),
                                         ("arithmetic", 3)])
def test_invalid_configuration_is_rejected_before_device_use(field, value):
    cfg = bh.SolverConfig(shape=(16, 16, 16), omega_plus=1.25, omega_minus=0.75)
    c = cfg.to_c()
    setattr(c, field, value)
    handle = ctypes.c_void_p()
    st = _abi.load().bf_create(ctypes.byref(c), ctypes.byref(handle))
    assert st == _abi.ERROR_INVALID_CONFIGURATION
    assert not handle.value


def test_strict_arithmetic_requires_fp32_storage():
    cfg = bh.SolverConfig(shape=(16, 16, 16), omega_plus=1.25, omega_minus=0.75, storage=bh.STORAGE_FP16_OFFSET,
                          arithmetic=bh.ARITHMETIC_STRICT)
    handle = ctypes.c_void_p()
    assert _abi.load().bf_create(ctypes.byref(cfg.to_c()), ctypes.byref(handle)) == _abi.ERROR_INVALID_CONFIGURATION


def test_mixed_periodic_axis_is_rejected():
    cfg = bh.SolverConfig(shape=(16, 16, 16), omega_plus=1.25, omega_minus=0.75,
                          boundary=((bh.BC_PERIODIC, bh.BC_FARFIELD), (0, 0), (0, 0)))
    handle = ctypes.c_void_p()
    assert _abi.load().bf_create(ctypes.byref(cfg.to_c()), ctypes.byref(handle)) == _abi.ERROR_INVALID_CONFIGURATION


def test_null_context_calls_fail_cleanly():
    lib = _abi.load()
    assert lib.bf_step(None, 1) == _abi.ERROR_INVALID_ARGUMENT
    assert lib.bf_step_index(None) == 0
    assert lib.bf_last_error(None) == b"null context"
    lib.bf_destroy(None)


def test_lattice_scaling_reproduces_reference_guards():
    # translating-sphere scaling: L = 8 cells, U = 0.05, Re = 4 -> nu 0.1, tau+ 0.8, magic 3/16 -> tau- 1.125
    s = bh.LatticeScaling.make(8.0, 8, 0.05, 0.05 * 8 / 0.1, 1.0, lattice_reference_speed=0.05)
    assert abs(s.tau_plus - 0.8) < 1e-6 and abs(s.tau_minus - 1.125) < 1e-6
    assert s.force_to_physical == 1.0 and s.velocity_to_lattice == 1.0
    with pytest.raises(bh.BirdFlowError):
        bh.LatticeScaling.make(8.0, 8, 1.0, 1e9, 1.0)          # tau+ below 0.50005
    with pytest.raises(bh.BirdFlowError):
        bh.LatticeScaling.make(8.0, 8, 1.0, 100.0, 1.0, lattice_reference_speed=0.2)   # Mach above 0.15


def test_create_without_gpu_reports_device_unavailable():
    name = ctypes.create_string_buffer(64)
    total = ctypes.c_uint64()
    if _abi.load().bf_device_info(name, 64, ctypes.byref(total)) == 0:
        pytest.skip("a device is present")
    with pytest.raises(bh.BirdFlowError) as e:
        bh.BirdFlowSimulation(bh.SolverConfig(shape=(16, 16, 16), omega_plus=1.25, omega_minus=0.75))
    assert e.value.status == _abi.ERROR_DEVICE_UNAVAILABLE
