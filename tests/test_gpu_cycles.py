"""Multi-cycle graphs (QCD_TUNE_CYCLES_PER_GRAPH): K consecutive cycles
recorded as one graph, each cycle's measurement (plaquette fused into the
last sweep, Polyakov loop, final sums) on a second stream beside the next
cycle's odd X, Y, Z half-steps, the odd T half-step waiting for it.  The
replays must give the same bits as single-cycle replays: state images
(links as stored, generator states, records and their counts) after whole
K-blocks plus single cycles, on the benchmarked workload and on lattices
with other storage, generators and group; a mask with the Wilson loop
(which reads every direction) keeps single-cycle graphs."""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case]", msg, flush=True)


CASES = [
    # (sun, dims, precision, storage, prng, n_or, cycles)
    (3, (32, 32, 32, 8), "fp32", 18, "XOR128", 4, 11),   # the benchmarked workload: one block of 8 + 3
    (3, (8, 16, 8, 4), "fp32", 12, "PM", 2, 19),
    (3, (6, 4, 4, 6), "fp64", 18, "RANLUX", 1, 10),       # PRN produce kernels inside the cycle
    (2, (8, 8, 8, 4), "fp64", 4, "XOR128", 3, 17),
]


def _run(case, k, mask_wilson=False):
    sun, dims, prec, storage, prng, n_or, cycles = case
    with Q.Lattice(sun=sun, dims=dims, beta=5.9 if sun == 3 else 2.3, precision=prec, storage_reals=storage,
                   prng=prng, randseries=31, max_measurements=cycles + 4) as lat:
        lat.tune(cycles_per_graph=k)
        if mask_wilson:
            lat.set_wilson(2, 2)
        lat.init("cold")
        lat.prng_init()
        lat.capture_cycle(n_hb=1, n_or=n_or, plaquette=True, polyakov=True, wilson=mask_wilson)
        captured = lat.tuning()["captured_cycles_per_graph"]
        lat.launch_cycle(cycles)
        lat.sync()
        plq, poly = lat.raw_measurements()
        assert len(plq) == cycles and len(poly) == cycles
        return lat.save_state(), captured


@pytest.mark.parametrize("case", CASES, ids=["su%d-%s-%s-%d-%s" % (c[0], "x".join(map(str, c[1])), c[2], c[3], c[4])
                                             for c in CASES])
def test_multi_cycle_graph_equals_single_cycles(case):
    _case("cycles per graph 8 vs 1: %s" % (case,))
    a, ka = _run(case, 8)
    b, kb = _run(case, 1)
    assert (ka, kb) == (8, 0)
    diff = int(np.count_nonzero(a != b))
    print("  differing bytes %d of %d" % (diff, a.size))
    assert a.size == b.size and diff == 0


def test_wilson_mask_keeps_single_cycles():
    _case("cycles per graph with the Wilson loop")
    case = CASES[1]
    a, ka = _run(case, 8, mask_wilson=True)
    b, kb = _run(case, 1, mask_wilson=True)
    assert (ka, kb) == (0, 0)
    assert int(np.count_nonzero(a != b)) == 0


@pytest.mark.parametrize("sun,dims,prec", [(3, (32, 32, 32, 8), "fp32"), (2, (8, 8, 8, 8), "fp64")],
                         ids=["su3-32x32x32x8-fp32", "su2-8x8x8x8-fp64"])
def test_heatbath_only_graph_equals_eager(sun, dims, prec):
    """The heatbath-only cycle with a plaquette record (the workload of
    bench.py --table) against the same sweeps and measurements launched one
    call at a time: links and generator states bit for bit.  The cycle
    sums the plaquette inside the sweep's even Y/Z/T half-steps, each trace
    formed in the links' precision, and the eager measurement with the
    standalone kernel, which forms the traces in FP64: the records agree to
    1e-8 relative for FP32 links (measured: 1.2e-10) and 1e-12 for FP64."""
    _case("heatbath-only cycle graph vs eager: SU(%d) %s %s" % (sun, dims, prec))
    out = []
    for graph in (True, False):
        with Q.Lattice(sun=sun, dims=dims, beta=5.9 if sun == 3 else 2.3, precision=prec, prng="XOR128",
                       randseries=54, max_measurements=12) as lat:
            lat.init("cold")
            lat.prng_init()
            if graph:
                lat.capture_cycle(n_hb=1, n_or=0, plaquette=True, polyakov=False)
                lat.launch_cycle(6)
            else:
                for _ in range(6):
                    lat.sweep("hb")
                    lat.measure(plaquette=True, polyakov=False)
            lat.sync()
            plq, poly = lat.raw_measurements()
            assert len(plq) == 6 and len(poly) == 0
            out.append((lat.export_links(), lat.prng_counter(), plq))
    (la, ca, pa), (lb, cb, pb) = out
    assert ca == cb
    assert la.shape == lb.shape and np.array_equal(la.view(np.uint8), lb.view(np.uint8))
    assert np.allclose(pa, pb, rtol=1e-8 if prec == "fp32" else 1e-12, atol=0.0)
