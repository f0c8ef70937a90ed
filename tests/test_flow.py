"""Learned SU(2) proposal on the CPU: the float64 restatement
(oracle/flow_ref.py) pinned by mathematics -- spline inverse and derivative,
the coupling layer's Jacobian with respect to Haar measure against finite
differences, gauge equivariance, invertibility -- and the FP32 PyTorch
training model checked against it."""
import math

import numpy as np
import pytest

import flow_ref as FR


def test_spline_inverse_derivative_and_identity():
    rng = np.random.default_rng(3)
    K = 4
    for trial in range(5):
        r = rng.normal(scale=1.5, size=3 * K + 1)
        xs = np.linspace(0.0, math.pi, 301)
        ys = [FR.spline(r, K, x)[0] for x in xs]
        assert abs(ys[0]) < 1e-15 and abs(ys[-1] - math.pi) < 1e-12
        assert np.all(np.diff(ys) > 0)
        for x in xs[1:-1:7]:
            y, dlog = FR.spline(r, K, x)
            xb, dlog_b = FR.spline(r, K, y, inverse=True)
            assert abs(xb - x) < 1e-10 and abs(dlog_b - dlog) < 1e-8
            h = 1e-6
            fd = (FR.spline(r, K, x + h)[0] - FR.spline(r, K, x - h)[0]) / (2 * h)
            assert abs(fd / math.exp(dlog) - 1) < 1e-5
    zero = np.zeros(3 * K + 1)
    for x in (0.1, 1.0, 2.5):
        y, dlog = FR.spline(zero, K, x)
        assert abs(y - x) < 1e-12 and abs(dlog) < 1e-12


def _qexp(e):
    n = np.linalg.norm(e)
    return np.concatenate([[math.cos(n)], (math.sin(n) / n) * e]) if n > 0 else np.array([1.0, 0, 0, 0])


def _qlog(q):
    n = np.linalg.norm(q[1:])
    return math.atan2(n, q[0]) * q[1:] / n


def test_coupling_jacobian_against_finite_differences():
    """det d(eps')/d(eps) at eps = 0 for U = exp(i eps.sigma) U0 -> U' =
    exp(i eps'.sigma) U0' equals exp(log J): log J is the density ratio with
    respect to Haar measure (flat in exponential coordinates at the origin)."""
    rng = np.random.default_rng(7)
    K = 4
    for trial in range(6):
        U0, Vq = FR.haar(rng, ()), FR.haar(rng, ())
        r = rng.normal(scale=1.0, size=3 * K + 1)
        U1, lj = FR.map_link(U0, Vq, r, K)
        J = np.zeros((3, 3))
        h = 1e-6
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            up = FR.map_link(FR.qmul(_qexp(e), U0), Vq, r, K)[0]
            dn = FR.map_link(FR.qmul(_qexp(-e), U0), Vq, r, K)[0]
            J[:, j] = (_qlog(FR.qmul(up, FR.qdag(U1))) - _qlog(FR.qmul(dn, FR.qdag(U1)))) / (2 * h)
        assert abs(math.log(abs(np.linalg.det(J))) - lj) < 1e-6
        # the inverse map undoes it and reports the same log J
        back, lj_b = FR.map_link(U1, Vq, r, K, inverse=True)
        assert np.max(np.abs(back - U0)) < 1e-12 and abs(lj_b - lj) < 1e-9


def _random_params(rng, layers, scale2=0.02):
    out = []
    for _ in range(layers):
        out.append((rng.normal(scale=0.1, size=(32, 648)), rng.normal(scale=0.1, size=32),
                    rng.normal(scale=scale2, size=(16, 2592)), rng.normal(scale=0.3, size=16)))
    return out


def test_layer_gauge_equivariance_and_inverse():
    rng = np.random.default_rng(11)
    lat = FR.Lat((4, 2, 2, 2))
    L = FR.haar(rng, (4, lat.V))
    prm = _random_params(rng, 8)
    g = FR.haar(rng, (lat.V,))
    Lg = np.empty_like(L)
    for mu in range(4):
        for s in range(lat.V):
            Lg[mu, s] = FR.qmul(FR.qmul(g[s], L[mu, s]), FR.qdag(g[lat.step(s, mu, 1)]))
    for l in (0, 1, 5):
        out, lj = FR.layer(lat, L, l, prm[l], 4)
        out_g, lj_g = FR.layer(lat, Lg, l, prm[l], 4)
        assert abs(lj - lj_g) < 1e-9
        for mu in range(4):
            for s in range(lat.V):
                want = FR.qmul(FR.qmul(g[s], out[mu, s]), FR.qdag(g[lat.step(s, mu, 1)]))
                assert np.max(np.abs(out_g[mu, s] - want)) < 1e-10
        back, lj_b = FR.layer(lat, out, l, prm[l], 4, inverse=True)
        assert np.max(np.abs(back - L)) < 1e-10 and abs(lj_b - lj) < 1e-8
        assert abs(lj) > 1e-3   # the layer is not the identity
    # the whole flow: forward then log q of the result
    Z = FR.haar(rng, (4, lat.V))
    U, lq = FR.forward(lat, Z, prm, 4)
    assert abs(FR.log_q(lat, U, prm, 4) - lq) < 1e-8


def test_torch_model_matches_oracle():
    torch = pytest.importorskip("torch")
    from qcdgpu_amd import flow as QF

    rng = np.random.default_rng(5)
    dims = (4, 2, 2, 2)
    lat = FR.Lat(dims)
    layers = 8
    m = QF.FlowModel(dims=dims, layers=layers, bins=4, device="cpu", seed=1)
    flat = m.flat_params()
    prm = _random_params(rng, layers)
    per = QF.LAYER_PARAMS
    for l, (W1, b1, W2, b2) in enumerate(prm):
        o = l * per
        flat[o:o + QF.W1N] = W1.ravel()
        flat[o + QF.W1N:o + QF.W1N + 32] = b1
        flat[o + QF.W1N + 32:o + QF.W1N + 32 + QF.W2N] = W2.ravel()
        flat[o + QF.W1N + 32 + QF.W2N:o + per] = b2
    with torch.no_grad():
        m.params.copy_(torch.as_tensor(flat))
    prm32 = FR.split_params(flat, layers)   # the float32 values, in float64
    U = FR.haar(rng, (3, 4, lat.V))
    lq_t = m.log_q(torch.as_tensor(U, dtype=torch.float32)).detach().numpy()
    S_t = m.action(torch.as_tensor(U, dtype=torch.float32), 2.2).detach().numpy()
    for b in range(3):
        lq_o = FR.log_q(lat, U[b], prm32, 4)
        assert abs(lq_t[b] - lq_o) < 2e-3 * max(1.0, abs(lq_o)), (lq_t[b], lq_o)
        assert abs(S_t[b] - FR.action(lat, U[b], 2.2)) < 1e-3
    Z = FR.haar(rng, (1, 4, lat.V))
    Uf, lq_f = m.forward(torch.as_tensor(Z, dtype=torch.float32))
    Uo, lq_o = FR.forward(lat, Z[0], prm32, 4)
    assert np.max(np.abs(Uf.detach().numpy()[0] - Uo)) < 1e-4
    assert abs(float(lq_f[0].detach()) - lq_o) < 2e-3 * max(1.0, abs(lq_o))


def test_quaternion_conventions():
    """The quaternion product is the 2x2 matrix product, and the canonical
    link conversion round-trips."""
    from qcdgpu_amd import flow as QF

    rng = np.random.default_rng(2)
    a, b = FR.haar(rng, ()), FR.haar(rng, ())
    assert np.allclose(FR.as_matrix(FR.qmul(a, b)), FR.as_matrix(a) @ FR.as_matrix(b), atol=1e-14)
    assert np.allclose(FR.as_matrix(FR.qdag(a)), FR.as_matrix(a).conj().T, atol=1e-14)
    q = FR.haar(rng, (4, 16)).astype(np.float32)
    can = QF.to_canonical(q)
    m = FR.as_matrix(q[2, 5].astype(np.float64))
    assert np.allclose(can[5, 2, ..., 0] + 1j * can[5, 2, ..., 1], m, atol=1e-7)
    assert np.array_equal(QF.to_quaternions(can, (2, 2, 2, 2)), q)
