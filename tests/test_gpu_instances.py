"""Generator streams in the reference's own layout for any PRNG_instances
(random.cpp:217-223: the reference takes the device's
CL_DEVICE_IMAGE3D_MAX_WIDTH): one produce() of PRNG_samples quads per
instance before every heatbath half-step (suncl.cpp:2302-2310), link i of
the half-step reading quad i + k PRNGSTEP of that buffer, i.e. instance
(i + k PRNGSTEP) mod I, sample floor((i + k PRNGSTEP) / I).  Compared with
the oracle on the same buffers: FP64 <= 1e-12 per link, FP32 <= 1e-5 per
link with no KP accept decision flipped.  Instance counts below, at and
above PRNGSTEP."""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case]", msg, flush=True)


def _oracle_streams(O, prng, seed, inst):
    if prng == "XOR128":
        return O.xor128_seed_table(seed, inst), O.xor128_produce
    if prng == "PM":
        return O.pm_seed_table(seed, inst), O.pm_produce
    nskip = [24, 48, 97, 223, 389][int(prng[-1])]
    return O.ranlux_seed_table(seed, inst), (lambda t, k: O.ranlux_produce(t, k, nskip))


def _oracle_hb(O, sun, dims, beta, prng, seed, inst, n_hb, start="cold"):
    L = O.lattice(sun, dims, beta, 10)
    V = int(np.prod(dims))
    prngstep = -(-(V // 2) // 64) * 64
    table, produce = _oracle_streams(O, prng, seed, inst)
    hot = start == "random"
    S = O.produce_quads(L, hot, inst)
    produce(table, S)   # the initial produce() of PRNG::initialize
    if hot:
        rowsize = -(-V // 64) * 64
        bufs = [produce(table, S).reshape(-1, 4) for _ in range(4)]
        ref = O.random_start(L, bufs, rowsize)
    else:
        ref = O.cold(L)
    for _ in range(n_hb):
        for parity in (1, 0):
            for mu in range(4):
                O.halfstep(L, ref, parity, mu, O.HB, O.QuadRng(produce(table, S), prngstep))
    return L, ref


CASES = [
    # sun, dims, prng, instances, precision, storage
    (3, (8, 4, 4, 4), "XOR128", 64, "fp64", 18),      # PRNGSTEP 256
    (3, (8, 4, 4, 4), "XOR128", 1024, "fp64", 18),
    (3, (6, 4, 4, 4), "PM", 128, "fp64", 18),         # V/2 = 192, PRNGSTEP 192
    (3, (8, 4, 4, 4), "RANLUX3", 192, "fp64", 18),
    (2, (8, 4, 4, 4), "XOR128", 320, "fp64", None),
    (2, (6, 4, 4, 4), "RANLUX2", 64, "fp64", None),
    (3, (8, 4, 4, 4), "XOR128", 2048, "fp32", 18),
    (3, (8, 4, 4, 4), "PM", 64, "fp32", 12),
]


@pytest.mark.parametrize("sun,dims,prng,inst,prec,storage", CASES,
                         ids=["su%d-%s-%s-I%d-%s-%s" % (c[0], "x".join(map(str, c[1])), c[2], c[3], c[4], c[5])
                              for c in CASES])
def test_heatbath_any_instances_vs_oracle(O, sun, dims, prng, inst, prec, storage):
    _case("instances %s" % ((sun, dims, prng, inst, prec, storage),))
    beta = 5.9 if sun == 3 else 2.3
    seed = 23
    L, ref = _oracle_hb(O, sun, dims, beta, prng, seed, inst, 2)
    gen = "RANLUX" if prng.startswith("RANLUX") else prng
    lux = int(prng[-1]) if gen == "RANLUX" else 3
    with Q.Lattice(sun=sun, dims=dims, beta=beta, precision=prec, storage_reals=storage, prng=gen, ranlux_luxury=lux,
                   randseries=seed, max_measurements=4, prng_instances=inst) as lat:
        lat.init("cold")
        lat.prng_init()
        lat.sweep("hb")
        lat.sweep("hb")
        got = lat.export_links()
        assert lat.prng_counter() == 17   # the initial produce() and one per heatbath half-step
    per_link = np.abs(got - ref).reshape(got.shape[0], 4, -1).max(axis=-1)
    if prec == "fp64":
        print("  max|dU| = %.3e" % per_link.max())
        assert per_link.max() <= 1e-12
    else:
        flips = int((per_link > 1e-3).sum())
        print("  flips=%d max|dU| = %.3e" % (flips, per_link.max()))
        assert flips == 0 and per_link.max() <= 1e-5


@pytest.mark.parametrize("sun,prng,inst", [(3, "XOR128", 64), (2, "PM", 320), (3, "RANLUX1", 128)])
def test_random_start_any_instances_vs_oracle(O, sun, prng, inst):
    """The reference's hot start (a produce() of the hot runs' size before
    each direction, suncl.cpp:5234-5248) from buffers of `inst` instances,
    then one heatbath sweep."""
    _case("random start instances %s" % ((sun, prng, inst),))
    dims, beta, seed = (8, 4, 4, 4), (5.9 if sun == 3 else 2.3), 41
    L, ref = _oracle_hb(O, sun, dims, beta, prng, seed, inst, 1, start="random")
    gen = "RANLUX" if prng.startswith("RANLUX") else prng
    lux = int(prng[-1]) if gen == "RANLUX" else 3
    with Q.Lattice(sun=sun, dims=dims, beta=beta, precision="fp64", prng=gen, ranlux_luxury=lux, randseries=seed,
                   max_measurements=4, prng_instances=inst) as lat:
        lat.init("random")
        lat.sweep("hb")
        got = lat.export_links()
    err = float(np.abs(got - ref).max())
    print("  max|dU| = %.3e" % err)
    assert err <= 1e-12


@pytest.mark.parametrize("prng", ["XOR128", "PM", "RANLUX"])
def test_instances_replay_and_state_image(prng):
    """Streams replayed to a production counter, and a state image loaded
    into a fresh lattice, continue exactly as the run itself."""
    _case("instances replay %s" % prng)
    kw = dict(sun=3, dims=(8, 4, 4, 4), beta=5.9, precision="fp64", prng=prng, randseries=5, max_measurements=4,
              prng_instances=128)
    with Q.Lattice(**kw) as a, Q.Lattice(**kw) as b, Q.Lattice(**kw) as c:
        a.init("cold")
        a.prng_init()
        for _ in range(3):
            a.sweep("hb")
        img = a.save_state()
        counter = a.prng_counter()
        b.init("cold")
        b.import_links(a.export_links())
        b.prng_replay(counter, "cold")
        c.load_state(img)
        for lat in (a, b, c):
            lat.sweep("hb")
            lat.sweep("or")
        ua = a.export_links()
        assert np.array_equal(ua, b.export_links())
        assert np.array_equal(ua, c.export_links())


def test_instances_refusals():
    _case("instances refusals")
    with Q.Lattice(sun=3, dims=(8, 4, 4, 4), beta=5.9, prng_instances=64, max_measurements=2) as lat:
        lat.prng_init()
        with pytest.raises(Q.QCDError, match="Haar"):
            lat.init("hot")
