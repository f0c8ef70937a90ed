"""Kernels by the reference's names (GPU::kernel_init / kernel_run with the
names of lattice_make_programs, suncl.cpp:3265-3648): a sweep run as the
reference's host runs it -- update_odd_X .. update_odd_T, update_even_X ..
update_even_T, then the plaquette and Polyakov kernels with their reduce
kernels -- gives the links and records of qcd_sweep + qcd_measure bit for
bit, also when recorded as a graph with qcd_graph_begin / qcd_graph_end and
replayed; start counts follow the runs; unknown names are refused."""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu

ORDER = ["update_odd_X", "update_odd_Y", "update_odd_Z", "update_odd_T",
         "update_even_X", "update_even_Y", "update_even_Z", "update_even_T"]
MEAS = ["lattice_measurement_plq", "reduce_measurement_plq_double2", "lattice_polyakov", "reduce_polyakov_double2"]


def _case(msg):
    print("[case]", msg, flush=True)


def _lat(prng):
    return Q.Lattice(sun=3, dims=(8, 4, 4, 4), beta=5.9, precision="fp64", prng=prng, randseries=3, max_measurements=8)


@pytest.mark.parametrize("prng", ["XOR128", "RANLUX"])
def test_named_kernels_equal_sweeps(prng):
    _case("registry %s" % prng)
    with _lat(prng) as a, _lat(prng) as b:
        for lat in (a, b):
            lat.init("cold")
            lat.prng_init()
        ids = [b.kernel(n) for n in ORDER + MEAS]
        for _ in range(2):
            a.sweep("hb")
            a.measure()
            for k in ids:
                b.run_kernel(k)
        assert np.array_equal(a.export_links(), b.export_links())
        ra, rb = a.raw_measurements(), b.raw_measurements()
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
        assert b.kernel_info(ids[0]) == ("update_odd_X", 2)
        # the same sequence recorded once as a graph and replayed twice
        b.graph_begin()
        for k in ids:
            b.run_kernel(k, wait=False)
        b.graph_end()
        b.launch_cycle(2)
        for _ in range(2):
            a.sweep("hb")
            a.measure()
        assert np.array_equal(a.export_links(), b.export_links())
        assert a.prng_counter() == b.prng_counter()
        ra, rb = a.raw_measurements(), b.raw_measurements()
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
        with pytest.raises(Q.QCDError, match="no kernel"):
            b.kernel("lattice_fourier_transform")
