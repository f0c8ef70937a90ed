"""Pins the oracle's per-site plaquette sums (orc_plaquette_sites), the
values the action per slice adds up (lattice_action_diff_x/y/z,
sun_measurements_cl.cl:186-615): against numpy products of the explicit
plaquette matrices site by site, their total against orc_plaquette (pinned
by the golden sweeps), and the cold-start closed form."""
import numpy as np


def _haar(rng, n, N):
    z = (rng.normal(size=(n, N, N)) + 1j * rng.normal(size=(n, N, N))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1.0 / N)


def test_plaquette_sites(O):
    dims = (3, 4, 2, 4)
    V = int(np.prod(dims))
    for sun in (3, 2):
        rng = np.random.default_rng(20 + sun)
        U = _haar(rng, 4 * V, sun).reshape(V, 4, sun, sun)
        links = np.ascontiguousarray(np.stack([U.real, U.imag], axis=-1))
        L = O.lattice(sun, dims)
        got = O.plaquette_sites(L, links)
        n1, n2, n3, n4 = dims
        x, t, z, y = np.indices((n1, n4, n3, n2)).reshape(4, -1)
        c = np.stack([x, y, z, t], axis=1)

        def shift(d):
            cc = c.copy()
            cc[:, d] = (cc[:, d] + 1) % dims[d]
            return cc[:, 1] + n2 * (cc[:, 2] + n3 * (cc[:, 3] + n4 * cc[:, 0]))

        want = np.zeros((V, 2))
        for a, b in [(0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3)]:
            P = U[:, a] @ U[shift(a), b] @ U[shift(b), a].conj().transpose(0, 2, 1) @ U[:, b].conj().transpose(0, 2, 1)
            want[:, 1 if b == 3 else 0] += np.trace(P, axis1=1, axis2=2).real
        assert np.max(np.abs(got - want)) < 1e-12
        assert np.allclose(got.sum(axis=0), O.plaquette(L, links), rtol=1e-13, atol=0)
        sl = O.action_slices(L, links)
        for k, n in zip("xyz", dims[:3]):
            assert sl[k].shape == (n, 2)
            assert np.allclose(sl[k].sum(axis=0), got.sum(axis=0), rtol=1e-13, atol=0)
        cold = O.action_slices(L, O.cold(L))
        for k, n in zip("xyz", dims[:3]):
            assert np.allclose(cold[k], 3 * sun * V // n)
