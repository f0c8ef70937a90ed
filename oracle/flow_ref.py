"""Restatement of the learned SU(2) proposal's mathematics in float64 numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
training model and the device flow, never on the product path.  The flow has
no counterpart in the reference; what is restated here is its definition:

  links       unit quaternions (a0, a1, a2, a3), U = a0 + i a.sigma,
              L[mu, site, :], site = y + z N2 + t N2 N3 + x N2 N3 N4
  layer l     mu = l % 4, mask parity m = (l // 4) % 2; active sites have
              sum_{nu != mu} x_nu + m even
  staple      A = sum_nu U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
                         + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
  map         V = A / |A|, W = U V = exp(i theta n.sigma), theta' = f(theta),
              U' = exp(i theta' n.sigma) V^+,
              log J = log f'(theta) + 2 log(sin theta' / sin theta)
  spline      rational-quadratic on [0, pi] (Durkan et al. 2019), K bins
  features    [|A| / 6 (active), plaquettes _|_ mu (3), mask, 0, 0, 0]
  conditioner conv 3^4 (8 -> 32) + b, tanh, conv 3^4 (32 -> 16) + b
"""
import math

import numpy as np

MIN_W = MIN_D = 1e-3


def qmul(a, b):
    a0, a1, a2, a3 = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    b0, b1, b2, b3 = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                     a0 * b1 + b0 * a1 - (a2 * b3 - a3 * b2),
                     a0 * b2 + b0 * a2 - (a3 * b1 - a1 * b3),
                     a0 * b3 + b0 * a3 - (a1 * b2 - a2 * b1)], axis=-1)


def qdag(a):
    return a * np.array([1.0, -1.0, -1.0, -1.0])


def as_matrix(q):
    """2x2 complex matrix of a quaternion (for cross-checks)."""
    a0, a1, a2, a3 = q
    return np.array([[a0 + 1j * a3, a2 + 1j * a1], [-a2 + 1j * a1, a0 - 1j * a3]])


class Lat:
    def __init__(self, dims):
        self.dims = tuple(dims)
        self.V = int(np.prod(dims))

    def coords(self, s):
        n1, n2, n3, n4 = self.dims
        y = s % n2
        r = s // n2
        z = r % n3
        r //= n3
        t = r % n4
        return [r // n4, y, z, t]

    def index(self, c):
        n1, n2, n3, n4 = self.dims
        x, y, z, t = (c[k] % self.dims[k] for k in range(4))
        return y + n2 * (z + n3 * (t + n4 * x))

    def step(self, s, d, o):
        c = self.coords(s)
        c[d] += o
        return self.index(c)

    def active(self, s, mu, par):
        c = self.coords(s)
        return (par + sum(c[d] for d in range(4) if d != mu)) % 2 == 0


def staple(lat, L, mu, s):
    A = np.zeros(4)
    xp = lat.step(s, mu, 1)
    for nu in range(4):
        if nu == mu:
            continue
        xpn, xmn = lat.step(s, nu, 1), lat.step(s, nu, -1)
        xpmn = lat.step(xmn, mu, 1)
        A += qmul(qmul(L[nu, xp], qdag(L[mu, xpn])), qdag(L[nu, s]))
        A += qmul(qmul(qdag(L[nu, xpmn]), qdag(L[mu, xmn])), L[nu, xmn])
    return A


def plaq0(lat, L, nu, rho, s):
    p = qmul(qmul(L[nu, s], L[rho, lat.step(s, nu, 1)]), qmul(qdag(L[nu, lat.step(s, rho, 1)]), qdag(L[rho, s])))
    return p[0]


def features(lat, L, mu, par):
    F = np.zeros((lat.V, 8))
    for s in range(lat.V):
        if lat.active(s, mu, par):
            F[s, 0] = np.linalg.norm(staple(lat, L, mu, s)) / 6.0
            F[s, 4] = 1.0
        k = 1
        for nu in range(4):
            for rho in range(nu + 1, 4):
                if mu not in (nu, rho):
                    F[s, k] = plaq0(lat, L, nu, rho, s)
                    k += 1
    return F


def conv(lat, X, W, b):
    """X (V, C), W (Cout, 81 C) with k = st * C + c, st = ((ox*3+oy)*3+oz)*3+ot."""
    V, C = X.shape
    out = np.tile(b, (V, 1)).astype(np.float64)
    for s in range(V):
        c0 = lat.coords(s)
        st = 0
        for ox in range(3):
            for oy in range(3):
                for oz in range(3):
                    for ot in range(3):
                        nb = lat.index([c0[0] + ox - 1, c0[1] + oy - 1, c0[2] + oz - 1, c0[3] + ot - 1])
                        out[s] += W[:, st * C:(st + 1) * C] @ X[nb]
                        st += 1
    return out


def spline_knots(r, K):
    def soft(v):
        e = np.exp(v - np.max(v))
        return e / e.sum()

    w = math.pi * (MIN_W + (1 - K * MIN_W) * soft(r[:K]))
    h = math.pi * (MIN_W + (1 - K * MIN_W) * soft(r[K:2 * K]))
    d = MIN_D + (1 - MIN_D) * np.logaddexp(0.0, r[2 * K:3 * K + 1]) / math.log(2.0)
    xk = np.concatenate([[0.0], np.cumsum(w)])
    yk = np.concatenate([[0.0], np.cumsum(h)])
    xk[-1] = yk[-1] = math.pi
    return xk, yk, d


def spline(r, K, v, inverse=False):
    """Monotone rational-quadratic map of [0, pi] onto itself; returns
    (image, log f'(x)) with x the preimage."""
    xk, yk, d = spline_knots(r, K)
    knots = yk if inverse else xk
    k = 0
    while k < K - 1 and v >= knots[k + 1]:
        k += 1
    w, h = xk[k + 1] - xk[k], yk[k + 1] - yk[k]
    s = h / w
    if not inverse:
        xi = min(max((v - xk[k]) / w, 0.0), 1.0)
    else:
        dy, dd = v - yk[k], d[k + 1] + d[k] - 2 * s
        a, bq, cq = h * (s - d[k]) + dy * dd, h * d[k] - dy * dd, -s * dy
        xi = min(max(2 * cq / (-bq - math.sqrt(max(bq * bq - 4 * a * cq, 0.0))), 0.0), 1.0)
    om = 1 - xi
    den = s + (d[k + 1] + d[k] - 2 * s) * xi * om
    num = s * s * (d[k + 1] * xi * xi + 2 * s * xi * om + d[k] * om * om)
    dlog = math.log(num) - 2 * math.log(den)
    if inverse:
        return xk[k] + xi * w, dlog
    return yk[k] + h * (s * xi * xi + d[k] * xi * om) / den, dlog


def map_link(U, Vq, r, K, inverse=False):
    """One active link: (U', log J of the forward map)."""
    W = qmul(U, Vq)
    sn = np.linalg.norm(W[1:])
    th = math.atan2(sn, W[0])
    th2, dlog = spline(r, K, th, inverse)
    th_in, th_out = (th2, th) if inverse else (th, th2)
    lj = dlog + 2 * math.log(math.sin(th_out) / math.sin(th_in))
    Wn = np.concatenate([[math.cos(th2)], W[1:] * (math.sin(th2) / sn)])
    return qmul(Wn, qdag(Vq)), lj


def layer(lat, L, l, params, K, inverse=False):
    """Coupling layer l with conditioner params (W1, b1, W2, b2)."""
    mu, par = l % 4, (l // 4) % 2
    W1, b1, W2, b2 = params
    F = features(lat, L, mu, par)
    r = conv(lat, np.tanh(conv(lat, F, W1, b1)), W2, b2)
    out = L.copy()
    lj = 0.0
    for s in range(lat.V):
        if lat.active(s, mu, par):
            A = staple(lat, L, mu, s)
            out[mu, s], j = map_link(L[mu, s], A / np.linalg.norm(A), r[s], K, inverse)
            lj += j
    return out, lj


def log_q(lat, L, layer_params, K):
    lj = 0.0
    for l in reversed(range(len(layer_params))):
        L, j = layer(lat, L, l, layer_params[l], K, inverse=True)
        lj += j
    return -lj


def forward(lat, Z, layer_params, K):
    lj = 0.0
    for l in range(len(layer_params)):
        Z, j = layer(lat, Z, l, layer_params[l], K)
        lj += j
    return Z, -lj


def action(lat, L, beta):
    return beta * sum(1.0 - plaq0(lat, L, nu, rho, s) for s in range(lat.V) for nu in range(4) for rho in range(nu + 1, 4))


def haar(rng, shape):
    q = rng.normal(size=shape + (4,))
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def split_params(flat, layers):
    """Flat device-layout vector -> per-layer (W1, b1, W2, b2)."""
    n1, n2 = 32 * 81 * 8, 16 * 81 * 32
    per = n1 + 32 + n2 + 16
    out = []
    for l in range(layers):
        p = np.asarray(flat[l * per:(l + 1) * per], dtype=np.float64)
        out.append((p[:n1].reshape(32, 648), p[n1:n1 + 32], p[n1 + 32:n1 + 32 + n2].reshape(16, 2592), p[n1 + 32 + n2:]))
    return out
