"""The C-ABI library loads and exports exactly the symbols declared in
include/hmc_engine.h; error paths forced from the host (no GPU work)."""
import os
import re

import pytest

from hmclab_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "hmc_engine.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(hmc_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    names = _declared()
    assert len(names) >= 25
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_device_count():
    assert _lib.lib().hmc_version() >= 100
    assert _lib.device_count() >= 0


def test_fault_injection_maps_to_exception(monkeypatch):
    monkeypatch.setenv("HMC_FAULT_INJECT", "hmc_device_count")
    with pytest.raises(_lib.InjectedFault, match="fault injected"):
        _lib.device_count()


def test_null_arguments_are_rejected():
    import ctypes as C

    with pytest.raises(ValueError):
        _lib.call("hmc_device_count", None)
    with pytest.raises(ValueError):
        _lib.call("hmc_problem_create", 0, 0, C.byref(C.c_void_p()))


def test_no_device_is_a_clear_runtime_error():
    # decided at run time, not at collection, so that collecting this module
    # never loads the library
    if _lib.device_count() > 0:
        pytest.skip("checks the no-device error path")
    from hmclab_amd.engine import Problem

    with pytest.raises(RuntimeError, match="no HIP device"):
        Problem(4)
