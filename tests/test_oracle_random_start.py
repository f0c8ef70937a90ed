"""Pins the oracle's restatement of the reference's hot start
(pyoracle.di_pierro, random_start, produce_quads; lattice_init_hot_X..T,
suncl.cl:56-161, lattice_random2 su2_update_cl.cl:110-146, lattice_random3
su3_update_cl.cl:165-225, PRNG_samples suncl.cpp:2302-2307).

(1) The SU(2) element against an independent construction: the matrix
exponential exp(i alpha n.sigma) of the Pauli form with the polar axis
n = (sin theta cos phi, sin theta sin phi, cos theta), built from numpy's
eigendecomposition, so a wrong angle scale, sign or component fails; its
trace is 2 cos(pi u) whatever v and w are.
(2) SU(3): the entries of E12(A) E13(B) E23(C) that the order and the
embeddings fix -- M11 = A11 B11, M21 = A21 B11, M31 = B21,
M12 = A12 C11 + A11 B12 C21 -- and the reduction to E12(A) when B = C = 1
(u = 0 gives the unit element); unitarity and det = 1.
(3) Addressing: changing one quad of one direction's buffer changes exactly
the link that reads it (site q mod rowsize for SU(3), q for SU(2)).
(4) The produce() size against the reference's own numbers: 33 quads per
instance for SU(3), NHIT 10, cold; 41 in a hot run (V/2 a multiple of 64)."""
import numpy as np
import pytest

PAULI = [np.array([[0, 1], [1, 0]], dtype=complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]],
                                                                                          dtype=complex)]


def _expi(H):
    w, v = np.linalg.eigh(H)
    return (v * np.exp(1j * w)) @ v.conj().T


def test_su2_element_is_the_polar_exponential(O):
    rng = np.random.default_rng(5)
    u, v, w = rng.random((3, 200))
    W = O.di_pierro(u, v, w)
    for k in range(200):
        alpha, phi, c = np.pi * u[k], 2 * np.pi * v[k], 2 * w[k] - 1
        s = np.sqrt(1 - c * c)
        n = (s * np.cos(phi), s * np.sin(phi), c)
        want = _expi(alpha * sum(n[i] * PAULI[i] for i in range(3)))
        assert np.max(np.abs(W[k] - want)) < 1e-13
        assert abs(np.trace(W[k]) - 2 * np.cos(np.pi * u[k])) < 1e-14
        assert abs(np.linalg.det(W[k]) - 1) < 1e-14


def _flat(rng, V, rowsize, k=3):
    return rng.random((k * rowsize + 64, 4)).astype(np.float32)


def test_su3_product_structure(O):
    dims = (4, 2, 4, 2)
    V = int(np.prod(dims))
    rowsize = -(-V // 64) * 64
    L = O.lattice(3, dims)
    rng = np.random.default_rng(8)
    bufs = [_flat(rng, V, rowsize) for _ in range(4)]
    links = O.random_start(L, bufs, rowsize)
    U = links[..., 0] + 1j * links[..., 1]
    for d in range(4):
        b = bufs[d].astype(np.float64)
        for s in range(V):
            q0, q1, q2 = b[s], b[s + rowsize], b[s + 2 * rowsize]
            A, B, C = (O.di_pierro(q0[c], q1[c], q2[c]) for c in range(3))
            M = U[s, d]
            assert abs(M[0, 0] - A[0, 0] * B[0, 0]) < 1e-14
            assert abs(M[1, 0] - A[1, 0] * B[0, 0]) < 1e-14
            assert abs(M[2, 0] - B[1, 0]) < 1e-14
            assert abs(M[0, 1] - (A[0, 1] * C[0, 0] + A[0, 0] * B[0, 1] * C[1, 0])) < 1e-14
            assert np.max(np.abs(M @ M.conj().T - np.eye(3))) < 1e-14
            assert abs(np.linalg.det(M) - 1) < 1e-14
    # B = C = 1 (their u = 0): the SU(3) link is E12(A)
    for b in bufs:
        b[:, 1:3] = np.where(np.arange(b.shape[0])[:, None] < rowsize, 0.0, b[:, 1:3])
    U = O.random_start(L, bufs, rowsize)
    U = U[..., 0] + 1j * U[..., 1]
    b = bufs[2].astype(np.float64)
    for s in range(V):
        A = O.di_pierro(b[s, 0], b[s + rowsize, 0], b[s + 2 * rowsize, 0])
        want = np.eye(3, dtype=complex)
        want[:2, :2] = A
        assert np.max(np.abs(U[s, 2] - want)) < 1e-14


@pytest.mark.parametrize("sun", [3, 2])
def test_addressing(O, sun):
    dims = (4, 4, 2, 2)
    V = int(np.prod(dims))
    rowsize = -(-V // 64) * 64
    L = O.lattice(sun, dims)
    rng = np.random.default_rng(13)
    bufs = [_flat(rng, V, rowsize) for _ in range(4)]
    base = O.random_start(L, bufs, rowsize)
    probes = [(1, 5), (3, 0), (2, V - 1)] + ([(0, rowsize + 7), (1, 2 * rowsize + 3)] if sun == 3 else [])
    for d, q in probes:
        mod = [b.copy() for b in bufs]
        mod[d][q] = rng.random(4).astype(np.float32)
        diff = np.abs(O.random_start(L, mod, rowsize) - base).reshape(V, 4, -1).max(axis=-1) > 0
        assert diff.sum() == 1 and diff[q % rowsize, d], (d, q)
    # quads past the ones the links read change nothing
    mod = [b.copy() for b in bufs]
    for b in mod:
        b[(3 if sun == 3 else 1) * rowsize:] = 0.5
        if sun == 2:
            b[V:] = 0.5
    assert np.array_equal(O.random_start(L, mod, rowsize), base)


def test_produce_size(O):
    cold = O.lattice(3, (32, 32, 32, 8), 5.9, 10)
    assert O.produce_quads(cold, False) == 33       # 3 (NHIT + 1) quads per link
    assert O.produce_quads(cold, True) == 41        # + ROWSIZE * ND / PRNGSTEP = 8
    su2 = O.lattice(2, (8, 8, 8, 8), 2.3, 10)
    assert O.produce_quads(su2, True) == 41
    odd = O.lattice(3, (6, 6, 6, 4), 5.9, 10)       # V/2 = 432: PRNGSTEP 448, ROWSIZE 896
    assert O.produce_quads(odd, False) == 33 and O.produce_quads(odd, True) == 1 + (4 * 896 + 33 * 448 - 1) // 448
    # other PRNG_instances (random.cpp:219-223: PRNG_samples = 1 + (samples - 1) / instances):
    # 4,325,376 quads in all for 32^3 x 8
    assert O.produce_quads(cold, False, 2048) == 2112 and O.produce_quads(cold, False, 16384) == 264
    assert O.produce_quads(cold, True, 2048) == 1 + (4 * 262144 + 4325376 - 1) // 2048
