"""Pins the oracle's twisted boundary (pyoracle.twist_matrix, twisted_view,
halfstep_twisted; the reference's TBC reads, su3_matrix_memory.cl:57-98,
su2_matrix_memory.cl:56-88).

(1) The diagonal matrix T against the reference's own arithmetic on the
stored planes (cos / sin factor vectors and the zwxy swizzle for SU(3), the
plane product matrix_times2 with Omega for SU(2)), carried out here on Haar
links and rebuilt into matrices: a wrong angle, sign or row fails.
(2) Closed forms on a cold lattice: the twist touches exactly the (x,y)
plaquettes whose Y links sit at x = N1 - 1 (two per site of that slice),
each then having ReTr T^+ (or ReTr T); every other plaquette, the temporal
ones included, stays N.
(3) The twisted half-step equals the plain one at phi = omega = 0, leaves
every other (parity, direction) untouched, and stores the twisted link's
update without removing the twist."""
import numpy as np
import pytest


def _haar(rng, n, N):
    z = (rng.normal(size=(n, N, N)) + 1j * rng.normal(size=(n, N, N))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1.0 / N)


def _ref_twist_su3(U, phi, omega):
    """lattice_table_3's TBC branch on the stored planes, then the matrix
    rebuilt with row 3 = conj(row 1 x row 2) (lattice_reconstruct3)."""
    uv1 = np.array([U[0, 0].real, U[0, 1].real, U[0, 2].real, U[1, 2].real])
    uv2 = np.array([U[0, 0].imag, U[0, 1].imag, U[0, 2].imag, U[1, 2].imag])
    uv3 = np.array([U[1, 0].real, U[1, 1].real, U[1, 0].imag, U[1, 1].imag])
    A, B = (phi + omega) / 2, (omega - phi) / 2
    a1 = np.array([np.cos(A)] * 3 + [np.cos(B)])
    a2 = np.array([np.sin(A)] * 3 + [np.sin(B)])
    a3 = np.full(4, np.cos(B))
    a4 = np.array([-np.sin(B), -np.sin(B), np.sin(B), np.sin(B)])
    n1 = uv1 * a1 - uv2 * a2
    n2 = uv2 * a1 + uv1 * a2
    n3 = uv3 * a3 + uv3[[2, 3, 0, 1]] * a4            # .zwxy
    r1 = n1[:3] + 1j * n2[:3]
    r2 = np.array([n3[0] + 1j * n3[2], n3[1] + 1j * n3[3], n1[3] + 1j * n2[3]])
    return np.stack([r1, r2, np.conj(np.cross(r1, r2))])


def _ref_twist_su2(U, phi):
    """lattice_table_2's TBC branch: matrix_times2(Omega, m) on the planes
    (Re u11, Re u12, Im u11, Im u12), Omega = (cos phi/2, 0, sin phi/2, 0)."""
    u = np.array([np.cos(phi / 2), 0.0, np.sin(phi / 2), 0.0])
    v = np.array([U[0, 0].real, U[0, 1].real, U[0, 0].imag, U[0, 1].imag])
    x = -u[3] * v[3] + u[0] * v[0] - u[1] * v[1] - u[2] * v[2]
    z = u[1] * v[3] + u[2] * v[0] - u[3] * v[1] + u[0] * v[2]
    y = -u[2] * v[3] + u[1] * v[0] + u[0] * v[1] + u[3] * v[2]
    w = u[0] * v[3] + u[3] * v[0] + u[2] * v[1] - u[1] * v[2]
    a, b = x + 1j * z, y + 1j * w
    return np.array([[a, b], [-np.conj(b), np.conj(a)]])


@pytest.mark.parametrize("sun", [3, 2])
def test_twist_matrix_vs_reference_plane_arithmetic(O, sun):
    rng = np.random.default_rng(3)
    for U in _haar(rng, 20, sun):
        phi, omega = rng.uniform(-np.pi, np.pi, size=2)
        T = O.twist_matrix(sun, phi, omega)
        want = _ref_twist_su3(U, phi, omega) if sun == 3 else _ref_twist_su2(U, phi)
        assert np.max(np.abs(T @ U - want)) < 1e-13
        assert abs(np.linalg.det(T) - 1) < 1e-14


@pytest.mark.parametrize("sun,phi,omega", [(3, 3.0, 0.0), (3, 0.7, -1.3), (3, 0.0, 2.0), (2, 3.0, 0.0), (2, 1.1, 5.0)])
def test_cold_plaquette_closed_form(O, sun, phi, omega):
    dims = (5, 4, 2, 4)
    V = int(np.prod(dims))
    L = O.lattice(sun, dims)
    view = O.twisted_view(L, O.cold(L), phi, omega)
    plq = O.plaquette(L, view)
    T = O.twist_matrix(sun, phi, omega)
    retr = np.trace(T).real                          # ReTr T = ReTr T^+
    n_twisted = 2 * V // dims[0]                     # (x,y) plaquettes at x = N1 - 2 and N1 - 1
    assert abs(plq[0] - (3 * V * sun - n_twisted * (sun - retr))) < 1e-10
    assert abs(plq[1] - 3 * V * sun) < 1e-10          # temporal planes: T and T^+ cancel


def test_twisted_halfstep(O):
    dims = (4, 4, 2, 4)
    V = int(np.prod(dims))
    L = O.lattice(3, dims, 5.9, 10)
    base = O.cold(L)
    r = O.SeqRng("RANLUX", 54, 223)
    O.sweep(L, base, O.HB, r)                         # a non-trivial configuration
    # phi = omega = 0: identical to the plain half-step
    a, b = base.copy(), base.copy()
    O.halfstep(L, a, 1, 1, O.HB, O.SeqRng("RANLUX", 9, 223))
    O.halfstep_twisted(L, b, 1, 1, O.HB, O.SeqRng("RANLUX", 9, 223), 0.0, 0.0)
    assert np.array_equal(a, b)
    # with a twist: only (parity 1, Y) links change; the twisted view's
    # update is what is stored
    phi, omega = 1.3, -0.4
    c = base.copy()
    O.halfstep_twisted(L, c, 1, 1, O.HB, O.SeqRng("RANLUX", 9, 223), phi, omega)
    sites = np.array([O.parity_site(L, i, 1) for i in range(V // 2)])
    mask = np.zeros((V, 4), bool)
    mask[sites, 1] = True
    assert np.array_equal(c[~mask], base[~mask])
    view = O.twisted_view(L, base, phi, omega)
    O.halfstep(L, view, 1, 1, O.HB, O.SeqRng("RANLUX", 9, 223))
    assert np.array_equal(c[mask], view[mask])
    # the twist changes the updates of links whose staples hold twisted links
    assert not np.array_equal(c[mask], a[mask])
