"""The C-ABI library loads and exports every function include/qcd_hip.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os

import pytest

import qcdgpu_amd
from qcdgpu_amd import _abi


def test_library_built():
    assert os.path.exists(_abi.LIB_PATH), "run __graft_entry__.build() first"


def test_every_header_symbol_exported():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = _abi.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = set(_abi.header_symbols())
    bound = set(_abi._SIGNATURES) | {"qcd_last_error"}
    assert names == bound


def test_abi_version_and_error_path():
    lib = _abi.lib()
    assert lib.qcd_abi_version() == 1
    # argument validation happens before any device call
    p = _abi.Params()
    p.sun = 4
    h = ctypes.c_void_p()
    rc = lib.qcd_lattice_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == 1
    assert b"sun" in lib.qcd_last_error()


def test_odd_extent_rejected():
    lib = _abi.lib()
    p = _abi.Params()
    p.sun = 3
    for i, v in enumerate((4, 3, 4, 4)):
        p.n[i] = v
    p.storage_reals = 18
    p.nhit = 10
    p.nhitpar = 1
    p.max_measurements = 4
    h = ctypes.c_void_p()
    assert lib.qcd_lattice_create(ctypes.byref(p), ctypes.byref(h)) == 1
    assert b"even" in lib.qcd_last_error()


def test_missing_library_raises(monkeypatch):
    monkeypatch.setattr(_abi, "_lib", None)
    monkeypatch.setattr(_abi, "LIB_PATH", "/nonexistent/libqcdhip.so")
    with pytest.raises(qcdgpu_amd.QCDError):
        _abi.lib()


@pytest.mark.parametrize("inst,stream,message", [(100, 0, b"multiple of 64"), (64, 1, b"reference stream layout")])
def test_prng_instances_validated(inst, stream, message):
    """PRNG_instances: 0 or a multiple of 64, reference stream layout only
    (checked before any device call)."""
    lib = _abi.lib()
    p = _abi.Params()
    p.sun = 3
    for i, v in enumerate((4, 4, 4, 4)):
        p.n[i] = v
    p.storage_reals = 18
    p.nhit = 10
    p.nhitpar = 1
    p.max_measurements = 4
    p.prng = _abi.PRNG_XOR128
    p.stream_mode = stream
    p.prng_instances = inst
    h = ctypes.c_void_p()
    assert lib.qcd_lattice_create(ctypes.byref(p), ctypes.byref(h)) == 1
    assert message in lib.qcd_last_error()
