"""Pins the oracle's F_munu measurement (orc_fmunu, the element formulas of
trace_lambda1..8 / trace_sigma1..3, su3_measurements_cl.cl:357-460,
su2_measurements_cl.cl:89-123).

Two independent routes: (1) a constant abelian flux b along one generator
in one plane, for which F = -i Tr(exp(i b G) G) / 2 = sin b at every site
and every other plane vanishes (a closed form that fixes the sign, the 1/2
and the plane order); (2) on Haar-random links, F from explicit Gell-Mann /
Pauli matrices and numpy traces of the full plaquette matrix, for every
generator, plane and both fields (catches a wrong element, sign or
transposition in any of the sixteen formulas)."""
import numpy as np
import pytest

GM = np.zeros((9, 3, 3), dtype=np.complex128)   # Gell-Mann lambda_1..8
GM[1][0, 1] = GM[1][1, 0] = 1
GM[2][0, 1], GM[2][1, 0] = -1j, 1j
GM[3][0, 0], GM[3][1, 1] = 1, -1
GM[4][0, 2] = GM[4][2, 0] = 1
GM[5][0, 2], GM[5][2, 0] = -1j, 1j
GM[6][1, 2] = GM[6][2, 1] = 1
GM[7][1, 2], GM[7][2, 1] = -1j, 1j
GM[8] = np.diag([1, 1, -2]) / np.sqrt(3)
PAULI = np.zeros((4, 2, 2), dtype=np.complex128)
PAULI[1] = [[0, 1], [1, 0]]
PAULI[2] = [[0, -1j], [1j, 0]]
PAULI[3] = [[1, 0], [0, -1]]

PLANES = {False: [(0, 1), (0, 2), (1, 2)], True: [(0, 3), (1, 3), (2, 3)]}


def _coords(dims):
    """(V, 4) coordinates in the oracle's site order (x slowest, then t, z, y)."""
    n1, n2, n3, n4 = dims
    x, t, z, y = np.indices((n1, n4, n3, n2)).reshape(4, -1)
    return np.stack([x, y, z, t], axis=1)


def _site(c, dims):
    n1, n2, n3, n4 = dims
    return c[:, 1] + n2 * (c[:, 2] + n3 * (c[:, 3] + n4 * c[:, 0]))


def _expi(b, G):
    w, v = np.linalg.eigh(G)
    return (v * np.exp(1j * b * w)) @ v.conj().T


def _to_links(U):
    return np.ascontiguousarray(np.stack([U.real, U.imag], axis=-1))


def _haar(rng, n, N):
    z = (rng.normal(size=(n, N, N)) + 1j * rng.normal(size=(n, N, N))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1.0 / N)


@pytest.mark.parametrize("sun,gen,efield", [(3, k, e) for k in range(1, 8) for e in (False, True)]
                         + [(2, k, e) for k in (1, 2, 3) for e in (False, True)])
def test_constant_flux_closed_form(O, sun, gen, efield):
    dims = (6, 4, 2, 4)
    V = int(np.prod(dims))
    c = _coords(dims)
    m = 1
    b = 2 * np.pi * m / dims[0]
    G = GM[gen] if sun == 3 else PAULI[gen]
    U = np.tile(np.eye(sun, dtype=np.complex128), (V, 4, 1, 1))
    bdir = 3 if efield else 1       # U_b(x) = exp(i b x1 G): flux b in the (x, b) plane
    for s in range(V):
        U[s, bdir] = _expi(b * c[s, 0], G)
    L = O.lattice(sun, dims)
    if sun == 3:
        g1, g2 = (gen, 8) if gen <= 4 else (3, gen)
        slot = 0 if gen <= 4 else 1
    else:
        g1, g2 = (gen, 2) if gen != 2 else (3, 2)
        slot = 0 if gen != 2 else 1
    F = O.fmunu(L, _to_links(U), efield, g1, g2)
    want = np.zeros(3, dtype=np.complex128)
    want[0] = V * np.sin(b)         # plane (x,y) or (x,t)
    assert np.allclose(F[slot], want, atol=1e-11, rtol=0)


@pytest.mark.parametrize("sun", [3, 2])
@pytest.mark.parametrize("efield", [False, True])
def test_random_links_vs_matrix_traces(O, sun, efield):
    dims = (4, 2, 4, 2)
    V = int(np.prod(dims))
    rng = np.random.default_rng(11 + sun)
    U = _haar(rng, V * 4, sun).reshape(V, 4, sun, sun)
    c = _coords(dims)
    L = O.lattice(sun, dims)

    def shift(d):
        cc = c.copy()
        cc[:, d] = (cc[:, d] + 1) % dims[d]
        return _site(cc, dims)

    sites = _site(c, dims)
    assert np.array_equal(sites, np.arange(V))
    gens = [(1, 5), (2, 6), (3, 7), (4, 8)] if sun == 3 else [(1, 2), (3, 2)]
    for g1, g2 in gens:
        F = O.fmunu(L, _to_links(U), efield, g1, g2)
        for pl, (a, b) in enumerate(PLANES[efield]):
            P = U[:, a] @ U[shift(a), b] @ U[shift(b), a].conj().transpose(0, 2, 1) @ U[:, b].conj().transpose(0, 2, 1)
            for slot, g in enumerate((g1, g2)):
                G = GM[g] if sun == 3 else PAULI[g]
                f = (-1j * np.trace(P @ G, axis1=1, axis2=2) / 2).sum()
                if sun == 2:
                    assert abs(f.imag) < 1e-12   # Tr(P sigma) is imaginary for P in SU(2)
                    f = f.real
                assert abs(F[slot, pl] - f) < 1e-12, (g, pl, F[slot, pl], f)
