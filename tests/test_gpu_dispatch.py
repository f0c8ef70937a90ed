"""The dispatch orders of the update kernels (which 64-link units each XCD
updates, and in which order) against the oracle, on lattices where the
benchmarked orders are the ones that run.

Modes 2 and 3 of QCD_TUNE_XCD_MAP are active only when N2/2 divides 64 and
the z-groups and units split evenly over the 8 XCDs; the oracle cases of
test_gpu_parity.py use lattices where they fall back to mode 1 or 0.  Here
`tuning()` confirms the mode before the comparison, so a remap error (a unit
updated twice, or never) shows up as a per-link difference against the
oracle.  Bars as elsewhere: FP64 <= 1e-12 per link; FP32 (18 and 12 reals)
<= 1e-5 per link with no KP accept decision flipped.
"""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case]", msg, flush=True)


_ORACLE = {}


def _oracle_run(O, dims, seed, n_hb, n_or):
    key = (dims, seed, n_hb, n_or)
    if key not in _ORACLE:
        _ORACLE[key] = _oracle_run_uncached(O, dims, seed, n_hb, n_or)
    L, ref = _ORACLE[key]
    return L, ref.copy()


def _oracle_run_uncached(O, dims, seed, n_hb, n_or):
    """The reference GPU path restated: XOR128 instances = PRNGSTEP, one
    produce() at seeding and one of S quads before every heatbath half-step
    (random.cpp:216-284, suncl.cpp:2302-2310); overrelaxation after."""
    L = O.lattice(3, dims, 5.9, 10)
    ref = O.cold(L)
    V = int(np.prod(dims))
    prngstep = -(-(V // 2) // 64) * 64
    S = 1 + (-(-(3 * prngstep * 11) // 64) * 64 - 1) // prngstep
    table = O.xor128_seed_table(seed, prngstep)
    O.xor128_produce(table, S)
    for _ in range(n_hb):
        for parity in (1, 0):
            for mu in range(4):
                O.halfstep(L, ref, parity, mu, O.HB, O.QuadRng(O.xor128_produce(table, S), prngstep))
    for _ in range(n_or):
        O.sweep(L, ref, O.OR, O.GidRng())
    return L, ref


# (dims, requested mode, expected (mode, XCD block x-values, z-rows))
CASES = [
    ((8, 16, 8, 4), 2, (2, 1, 8)),    # x-slabs of one x-value, the whole z extent
    ((8, 16, 8, 4), 3, (3, 1, 8)),    # one z-group: mode 3 keeps the slabs
    ((8, 16, 16, 4), 3, (3, 2, 8)),   # blocks of 2 x-values x 8 z-rows
    ((16, 16, 16, 4), 3, (3, 4, 8)),  # blocks of 4 x-values x 8 z-rows
    ((8, 32, 8, 4), 2, (2, 1, 8)),    # N2 = 32 as in the benchmark: 4 z-rows per unit
    ((8, 32, 8, 4), 3, (3, 2, 4)),
]
STORAGE = [("fp64", 18), ("fp32", 18), ("fp32", 12)]


@pytest.mark.parametrize("prec,storage", STORAGE, ids=["%s-%d" % s for s in STORAGE])
@pytest.mark.parametrize("dims,mode,expect", CASES, ids=["%s-m%d" % ("x".join(map(str, c[0])), c[1]) for c in CASES])
@pytest.mark.parametrize("shape", [3, 1])
def test_dispatch_order_vs_oracle(O, dims, mode, expect, prec, storage, shape):
    _case("dispatch %s mode %d %s-%d shape %d" % (dims, mode, prec, storage, shape))
    seed = 11
    L, ref = _oracle_run(O, dims, seed, 2, 1)
    with Q.Lattice(sun=3, dims=dims, beta=5.9, precision=prec, storage_reals=storage, randseries=seed,
                   max_measurements=4) as lat:
        lat.tune(xcd_map=mode, update_shape=shape)
        t = lat.tuning()
        assert (t["xcd_map"],) + tuple(t["xcd_block"]) == expect, t
        lat.init("cold")
        lat.prng_init()
        lat.sweep("hb")
        lat.sweep("hb")
        lat.sweep("or")
        got = lat.export_links()
    per_link = np.abs(got - ref).reshape(got.shape[0], 4, -1).max(axis=-1)
    if prec == "fp64":
        print("  max|dU| = %.3e" % per_link.max())
        assert per_link.max() <= 1e-12
    else:
        flips = int((per_link > 1e-3).sum())
        print("  flips=%d max|dU| = %.3e" % (flips, per_link.max()))
        assert flips == 0
        assert per_link.max() <= 1e-5


def _c1_run(mode, shape, n_cycles=2):
    dims = (32, 32, 32, 8)
    with Q.Lattice(sun=3, dims=dims, beta=5.9, precision="fp32", randseries=54, max_measurements=8) as lat:
        lat.tune(xcd_map=mode, update_shape=shape)
        t = lat.tuning()
        assert t["xcd_map"] == mode, t
        if mode == 2:
            assert t["xcd_block"] == (4, 32), t
        if mode == 3:
            assert t["xcd_block"] == (8, 16), t
        lat.init("cold")
        lat.prng_init()
        lat.capture_cycle(n_hb=1, n_or=4, plaquette=True, polyakov=True)
        lat.launch_cycle(n_cycles)
        lat.sync()
        plq, poly = lat.raw_measurements()
        assert len(plq) == n_cycles and len(poly) == n_cycles
        return lat.save_state(), lat.export_links()


def test_dispatch_orders_bit_identical_c1():
    """SU(3) 32^3 x 8 FP32 18-real (the benchmarked workload): two captured
    cycles of 1 heatbath + 4 overrelaxation sweeps with plaquette and
    Polyakov records under every dispatch order give byte-identical state
    images (links as stored, generator states, records and their counts)."""
    images = {}
    for mode in (0, 1, 2, 3):
        _case("C1 cycles xcd_map=%d shape=3" % mode)
        images[mode] = _c1_run(mode, 3)[0]
    base = images[0]
    for k, img in images.items():
        assert img.size == base.size
        diff = int(np.count_nonzero(img != base))
        print("  mode %d: %d differing bytes" % (k, diff))
        assert diff == 0


@pytest.mark.parametrize("n_cycles", [1, 2])
def test_update_shapes_identical_c1(n_cycles):
    """The one-wavefront kernel shape against the three-wavefront one at the
    benchmarked workload (both in dispatch order 3): state images equal byte
    for byte."""
    _case("C1 shapes, %d cycles" % n_cycles)
    img3, u3 = _c1_run(3, 3, n_cycles)
    img1, u1 = _c1_run(3, 1, n_cycles)
    d = np.abs(u1 - u3).reshape(u1.shape[0], 4, -1).max(axis=-1)
    nb = int(np.count_nonzero(img1 != img3))
    print("  differing bytes %d; links differing %d of %d, max|dU| %.3e" % (nb, int((d > 0).sum()), d.size, d.max()))
    assert nb == 0
