"""Race check of the update kernels: every half-step run in shadow mode
(staples gathered from a copy of the table taken just before the launch,
the updated links written into the table) against the same half-steps run
in place.  A (parity, mu) half-step reads only links of the other parity or
of other directions, so the two must give identical bits: state images
(links as stored, generator states, records) compared byte for byte.  A
kernel that read a link the same launch writes -- a wrong parity or
neighbour step, a unit updated by two workgroups -- makes them differ.
tests/test_oracle_shadow.py shows on the CPU that this comparison catches
such a read."""
import numpy as np
import pytest

import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case]", msg, flush=True)


CASES = [
    # (sun, dims, precision, storage, prng, instances, xcd_map, shape)
    (3, (8, 16, 8, 4), "fp32", 18, "XOR128", 0, 2, 3),   # the benchmarked order (mode 2), 3 wavefronts
    (3, (8, 32, 8, 4), "fp32", 18, "XOR128", 0, 3, 1),   # mode 3, one wavefront per 64 links
    (3, (6, 4, 4, 6), "fp64", 12, "RANLUX", 0, 1, 3),    # reference PRN buffer, 12-real storage
    (3, (8, 4, 4, 4), "fp32", 18, "XOR128", 128, 1, 3),  # PRNG_instances buffer layout
    (2, (8, 8, 8, 4), "fp64", 4, "PM", 0, 0, 3),         # SU(2) quaternions, hardware order
]


def _run(case, shadow):
    sun, dims, prec, storage, prng, inst, mode, shape = case
    with Q.Lattice(sun=sun, dims=dims, beta=5.9 if sun == 3 else 2.3, precision=prec, storage_reals=storage,
                   prng=prng, randseries=21, max_measurements=16, prng_instances=inst) as lat:
        lat.tune(xcd_map=mode, update_shape=shape, shadow=shadow)
        assert lat.tuning()["shadow"] == int(shadow)
        lat.init("cold")
        lat.prng_init()
        for _ in range(2):
            lat.sweep("hb")
            lat.sweep("or")
            lat.measure()
        # a captured cycle: the fused plaquette half-steps, replayed twice
        lat.capture_cycle(n_hb=1, n_or=2, plaquette=True, polyakov=True)
        lat.launch_cycle(2)
        lat.sync()
        return lat.save_state()


@pytest.mark.parametrize("case", CASES, ids=["su%d-%s-%s-%d-%s-i%d-m%d-s%d" % (c[0], "x".join(map(str, c[1])), c[2], c[3],
                                                                             c[4], c[5], c[6], c[7]) for c in CASES])
def test_shadow_equals_in_place(case):
    _case("shadow %s" % (case,))
    a = _run(case, False)
    b = _run(case, True)
    assert a.size == b.size
    diff = int(np.count_nonzero(a != b))
    print("  differing bytes %d of %d" % (diff, a.size))
    assert diff == 0
