"""GPU parity of the reference's hot start (QCD_START_RANDOM,
lattice_init_hot_X..T, suncl.cl:56-161) against the oracle on the
reference's produce() buffers (suncl.cpp:5234-5248), and of the heatbath
that follows it with the hot runs' produce() size (suncl.cpp:2304-2305).

Bars: FP64 links <= 1e-12, FP32 links <= 1e-5 (per real component; the
links are unitary, so this is the relative error); streams bit-exact
(replay and state images)."""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def _case(msg):
    print("[case]", msg, flush=True)


def _lat(sun, dims, prec, storage=None, prng="XOR128", seed=54, beta=5.9, stream="reference"):
    lux = 3
    if prng.startswith("RANLUX"):
        lux = int(prng[6:] or 3)
        prng = "RANLUX"
    return Q.Lattice(sun=sun, dims=dims, beta=beta, precision=prec, storage_reals=storage, nhit=10, prng=prng,
                     ranlux_luxury=lux, stream=stream, randseries=seed, max_measurements=16)


def _oracle_streams(O, L, prng, seed):
    """Seed table for PRNG_instances = PRNGSTEP and its produce()."""
    V = int(np.prod(list(L.n)))
    n = -(-(V // 2) // 64) * 64
    if prng == "XOR128":
        return O.xor128_seed_table(seed, n), O.xor128_produce
    if prng == "PM":
        return O.pm_seed_table(seed, n), O.pm_produce
    nskip = [24, 48, 97, 223, 389][int(prng[-1])]
    return O.ranlux_seed_table(seed, n), (lambda t, k: O.ranlux_produce(t, k, nskip))


def _oracle_random_run(O, L, prng, seed):
    """The reference's hot run: initial produce(), then one produce() before
    each direction's init kernel, then one before every heatbath half-step,
    all of S_hot quads per instance.  Returns the start and a function that
    applies the next heatbath sweep to given links."""
    V = int(np.prod(list(L.n)))
    n = -(-(V // 2) // 64) * 64
    rowsize = -(-V // 64) * 64
    S = O.produce_quads(L, True)
    table, produce = _oracle_streams(O, L, prng, seed)
    produce(table, S)
    bufs = [produce(table, S).reshape(-1, 4) for _ in range(4)]

    def sweep(links):
        for parity in (1, 0):
            for mu in range(4):
                O.halfstep(L, links, parity, mu, O.HB, O.QuadRng(produce(table, S), n))

    return O.random_start(L, bufs, rowsize), sweep


CASES = [
    (3, (6, 4, 2, 4), "fp64", 18, "XOR128", 54),
    (3, (4, 4, 4, 4), "fp64", 12, "PM", 7),
    (2, (6, 4, 2, 4), "fp64", 4, "RANLUX3", 1),
    (3, (6, 6, 6, 4), "fp64", 18, "RANLUX0", 5),     # V/2 = 432: idle instances, ROWSIZE != 2 PRNGSTEP
    (3, (6, 4, 2, 4), "fp32", 18, "XOR128", 54),
    (3, (4, 4, 4, 4), "fp32", 12, "PM", 9),
    (2, (4, 4, 4, 4), "fp32", 4, "XOR128", 3),
    (3, (8, 8, 8, 8), "fp64", 18, "XOR128", 11),
    (3, (6, 4, 2, 4), "fp32", 12, "RANLUX3", 21),
]
IDS = ["su%d-%s-%s-%d-%s" % (c[0], "x".join(map(str, c[1])), c[2], c[3], c[4]) for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_random_start_and_sweep_vs_oracle(O, case):
    sun, dims, prec, storage, prng, seed = case
    _case("random start %s" % (case,))
    L = O.lattice(sun, dims, 5.9 if sun == 3 else 2.3, 10)
    want_start, sweep = _oracle_random_run(O, L, prng, seed)
    with _lat(sun, dims, prec, storage, prng, seed, beta=L.beta) as lat:
        lat.init("random")
        assert lat.prng_counter() == 5
        got_start = lat.export_links()
        e0 = np.abs(got_start - want_start).max()
        lat.sweep("hb")
        assert lat.prng_counter() == 13
        got = lat.export_links()
    want = got_start.copy()   # one sweep on the configuration the device holds
    sweep(want)
    per_link = np.abs(got - want).reshape(got.shape[0], 4, -1).max(axis=-1)
    print("  start max|dU| = %.3e, after one sweep %.3e" % (e0, per_link.max()))
    assert e0 <= TOL[prec]
    if prec == "fp32":   # an accept decision within rounding of its threshold would flip: none expected here
        assert int((per_link > 1e-3).sum()) == 0
    assert per_link.max() <= TOL[prec]


def test_cold_start_after_random_start_uses_cold_size(O):
    """After a random start the other starts return to the cold runs'
    produce() size: a cold start then gives the cold run's streams."""
    _case("random then cold")
    dims = (6, 4, 2, 4)
    with _lat(3, dims, "fp64") as a, _lat(3, dims, "fp64") as b:
        a.init("random")
        a.sweep("hb")
        a.init("cold")
        a.prng_init()
        a.sweep("hb")
        b.init("cold")
        b.prng_init()
        b.sweep("hb")
        assert np.array_equal(a.export_links(), b.export_links())
        assert np.array_equal(a.save_state(), b.save_state())


def test_fast_streams_refused():
    with _lat(3, (4, 4, 4, 4), "fp32", stream="fast") as lat:
        with pytest.raises(Q.QCDError, match="QCD_STREAM_REFERENCE"):
            lat.init("random")


@pytest.mark.parametrize("prng", ["XOR128", "PM", "RANLUX2"])
def test_random_run_state_image_and_replay(prng):
    """A random-start run continues bit-identically from its state image on a
    fresh lattice (which starts at the cold size), and replaying its counter
    with start "random" reproduces its streams."""
    _case("random image/replay %s" % prng)
    dims = (6, 6, 6, 4)
    with _lat(3, dims, "fp32", 12, prng) as a, _lat(3, dims, "fp32", 12, prng) as b, \
            _lat(3, dims, "fp32", 12, prng) as c:
        a.init("random")
        a.sweep("hb")
        a.sweep("or")
        b.load_state(a.save_state())
        for x in (a, b):
            x.sweep("hb")
            x.halfstep(0, 2, "hb")
        assert np.array_equal(a.export_links(), b.export_links())
        assert np.array_equal(a.save_state(), b.save_state())
        cnt = a.prng_counter()
        assert cnt == 5 + 8 * 2 + 1
        c.prng_replay(cnt, "random")
        c.import_links(b.export_links())
        b.sweep("hb")
        c.sweep("hb")
        assert np.array_equal(b.export_links(), c.export_links())
