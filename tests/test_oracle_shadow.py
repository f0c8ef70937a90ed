"""The principle of the GPU race check (tests/test_gpu_shadow.py), shown on
the oracle: a half-step gathered from a copy of the links taken before it
equals the in-place half-step bit for bit, because a (parity, mu) half-step
reads only links it does not write; a pass that also updates the other
parity of the same direction (a planted read of links the same pass
writes) gives different links in the two forms, so the comparison catches
it."""
import numpy as np
import pytest


def _thermal(O, sun, dims):
    L = O.lattice(sun, dims, 5.9 if sun == 3 else 2.3, 10)
    links = O.cold(L)
    for _ in range(3):
        O.sweep(L, links, O.HB, O.GidRng())
    return L, links


@pytest.mark.parametrize("sun", [3, 2])
@pytest.mark.parametrize("mode", ["HB", "OR"])
def test_double_buffered_halfstep_equals_in_place(O, sun, mode):
    L, links = _thermal(O, sun, (4, 4, 2, 4))
    m = getattr(O, mode)
    for parity in (1, 0):
        for mu in range(4):
            a = links.copy()
            b = links.copy()
            O.halfstep(L, a, parity, mu, m, O.GidRng())
            O.halfstep_src(L, b, links.copy(), parity, mu, m, O.GidRng())
            assert np.array_equal(a, b)
            assert not np.array_equal(a, links)   # the half-step changed links
            links = a


@pytest.mark.parametrize("sun", [3, 2])
def test_planted_same_pass_read_is_caught(O, sun):
    L, links = _thermal(O, sun, (4, 4, 2, 4))
    for mu in range(4):
        a = links.copy()
        b = links.copy()
        O.halfstep_src(L, a, a, 2, mu, O.OR, O.GidRng())   # in place: reads what it wrote
        O.halfstep_src(L, b, links.copy(), 2, mu, O.OR, O.GidRng())          # from the copy
        d = np.abs(a - b).max()
        assert d > 1e-3, (mu, d)
