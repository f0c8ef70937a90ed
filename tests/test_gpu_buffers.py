"""Device buffers by id (qcd_buffer_*), the GPU::buffer_init / buffer_write /
buffer_map / buffer_kill registry of the reference's device layer
(clinterface.cpp:1296-1440,1558-1590): creation copies the host data for
the INPUT / IO / CONSTANT types and zeroes the others, _write uploads the
host copy again, _map returns the device contents, _unmap writes a changed
mirror back, LDS buffers hold no memory, killed ids are refused."""
import ctypes

import numpy as np
import pytest

from qcdgpu_amd import QCDError, _abi
from qcdgpu_amd._abi import check, lib

pytestmark = pytest.mark.gpu


def _create(kind, arr=None, count=None, elem=4):
    i = ctypes.c_int()
    ptr = arr.ctypes.data_as(ctypes.c_void_p) if arr is not None else None
    n = arr.size if arr is not None else count
    e = arr.itemsize if arr is not None else elem
    check(lib().qcd_buffer_create(kind, n, e, ptr, ctypes.byref(i)))
    return i.value


def _map(i, n, dtype):
    p = ctypes.c_void_p()
    check(lib().qcd_buffer_map(i, ctypes.byref(p)))
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_ubyte)), shape=(n * np.dtype(dtype).itemsize,)) \
        .view(dtype)


def test_buffer_registry_round_trips():
    print("[case] buffer registry", flush=True)
    check(lib().qcd_device_open(0, 0))
    host = np.arange(1000, dtype=np.float32)
    a = _create(_abi.BUFFER_IO, host)
    assert np.array_equal(_map(a, 1000, np.float32), host)
    host[:] = -host                       # buffer_write uploads the host copy again
    check(lib().qcd_buffer_write(a))
    assert np.array_equal(_map(a, 1000, np.float32), host)
    m = _map(a, 1000, np.float32)         # a changed mirror goes back with unmap
    m[10] = 123.0
    check(lib().qcd_buffer_unmap(a))
    host[:] = 0.0
    assert _map(a, 1000, np.float32)[10] == 123.0
    g = _create(_abi.BUFFER_GLOBAL, count=77, elem=8)
    assert not _map(g, 77, np.float64).any()
    nbytes, dev = ctypes.c_size_t(), ctypes.c_void_p()
    check(lib().qcd_buffer_info(g, ctypes.byref(nbytes), ctypes.byref(dev)))
    assert nbytes.value == 616 and dev.value
    lds = _create(_abi.BUFFER_LDS, count=256, elem=16)
    check(lib().qcd_buffer_info(lds, ctypes.byref(nbytes), ctypes.byref(dev)))
    assert nbytes.value == 4096 and not dev.value
    with pytest.raises(QCDError, match="LDS"):
        check(lib().qcd_buffer_write(lds))
    with pytest.raises(QCDError, match="host pointer"):
        _create(_abi.BUFFER_CONSTANT, count=4)
    for i in (a, g, lds):
        check(lib().qcd_buffer_kill(i))
    with pytest.raises(QCDError, match="no buffer"):
        check(lib().qcd_buffer_write(a))
    assert lib().qcd_buffer_size_align(1000, 64) == 1024 and lib().qcd_buffer_size_align(5, 0) == 1024
