"""The allocation-free measurement entry point (qcd_measure_noalloc) against
qcd_measure, through the C ABI: the same configuration measured by both must
give the same records bit for bit (every reduction runs in a fixed order).
Two calls in a row, the second on a smaller lattice, so the device array
holds the first call's partials when the second starts; and a lattice whose
buffers exceed the array is refused with an error, before any launch."""
import ctypes

import numpy as np
import pytest

import qcdgpu_amd as Q
from qcdgpu_amd import _abi

pytestmark = pytest.mark.gpu

ALL = (_abi.MEAS_PLAQUETTE | _abi.MEAS_POLYAKOV | _abi.MEAS_WILSON | _abi.MEAS_POLYAKOV_SLICES
       | _abi.MEAS_ACTION_SLICES)


def _case(msg):
    print("[case]", msg, flush=True)


def _records(lat):
    plq, poly = lat.raw_measurements()
    ps, acs = lat.polyakov_slices(raw=True), lat.action_slices(raw=True)
    return [plq, poly, lat.wilson_measurements(raw=True), np.concatenate([ps[k] for k in "xyz"], axis=1),
            np.concatenate([acs[k] for k in "xyz"], axis=1), lat.fmunu(raw=True)]


def _measure_both(dims, mask):
    """The same configuration measured by qcd_measure, then by
    qcd_measure_noalloc: records 0 and 1 of every kind."""
    lib = _abi.lib()
    with Q.Lattice(sun=3, dims=dims, beta=5.9, precision="fp64", randseries=31, max_measurements=8) as lat:
        lat.prng_init()
        lat.init("hot")
        lat.sweep("hb")
        lat.sweep("or")
        lat.sync()
        for fn in (lib.qcd_measure, lib.qcd_measure_noalloc):
            assert fn(lat._h, mask) == 0, lib.qcd_last_error()
        lat.sync()
        return _records(lat)


MASKS = [ALL | _abi.MEAS_FMUNU,                       # every kernel and the finaliser
         _abi.MEAS_PLAQUETTE | _abi.MEAS_POLYAKOV]    # final sums inside the Polyakov launch


@pytest.mark.parametrize("mask", MASKS, ids=["all", "plq-poly"])
def test_noalloc_measure_matches_measure(mask):
    for dims in ((8, 8, 8, 4), (6, 4, 4, 4)):   # the second call on a smaller lattice
        _case("noalloc measure %s mask %d" % (dims, mask))
        recs = _measure_both(dims, mask)
        assert len(recs[0]) == 2 and len(recs[1]) == 2   # plaquette and Polyakov records of both calls
        for r in recs:
            if len(r):
                assert len(r) == 2
                assert np.array_equal(r[0], r[1]), (dims, r)


def test_noalloc_refuses_oversized_lattice():
    _case("noalloc oversized")
    lib = _abi.lib()
    with Q.Lattice(sun=3, dims=(32, 32, 32, 32), beta=5.9, precision="fp32", randseries=1, max_measurements=2) as lat:
        lat.init("cold")
        rc = lib.qcd_measure_noalloc(lat._h, ALL)
        assert rc == 1
        assert b"too large for the allocation-free" in lib.qcd_last_error()
        # without the per-site buffers the same lattice fits
        assert lib.qcd_measure_noalloc(lat._h, _abi.MEAS_PLAQUETTE | _abi.MEAS_POLYAKOV) == 0, lib.qcd_last_error()
        lat.sync()
        plq, _ = lat.raw_measurements()
        assert np.allclose(np.array(plq)[0].sum() / (6 * 32 ** 4), 3.0, rtol=0, atol=1e-12)
