"""Learned global proposal for SU(2) lattices (qcd_flow_* in include/qcd_hip.h).

`Flow` drives the device flow (BF16 MFMA convolutions, spline coupling
layers) that draws proposals and evaluates log q.  `FlowModel` is the same
flow in FP32 PyTorch, used to train the parameters by forward KL on
configurations of the plain heatbath; its flat parameter vector has the
device layout.  `imh_chain` runs the heatbath chain with an
independence-Metropolis step after every `hb_per_step` heatbath sweeps:
accept U' with probability min(1, exp(-S(U') + S(U) + log q(U) - log q(U'))),
which keeps the Wilson-action distribution exact whatever the flow's
quality.  No counterpart in the reference (new work).
"""
import ctypes
import math
import time

import numpy as np

from . import _abi
from ._abi import check, lib

HIDDEN, CIN0, COUT = 32, 8, 16
STENCIL = 81
W1N, W2N = HIDDEN * STENCIL * CIN0, COUT * STENCIL * HIDDEN
LAYER_PARAMS = W1N + HIDDEN + W2N + COUT
MIN_W, MIN_D = 1e-3, 1e-3


# ---------------------------------------------------------------- link formats
def to_quaternions(links, dims):
    """Canonical links (V, 4, 2, 2, 2) -> float32 (4, V, 4) quaternions
    (a0, a1, a2, a3) with U = [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]]."""
    L = np.asarray(links).reshape(-1, 4, 2, 2, 2)
    q = np.stack([L[:, :, 0, 0, 0], L[:, :, 0, 1, 1], L[:, :, 0, 1, 0], L[:, :, 0, 0, 1]], axis=-1)
    return np.ascontiguousarray(q.transpose(1, 0, 2), dtype=np.float32)


def to_canonical(q):
    """(4, V, 4) quaternions -> canonical (V, 4, 2, 2, 2) float64 links."""
    q = np.asarray(q, dtype=np.float64).transpose(1, 0, 2)
    a0, a1, a2, a3 = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    out = np.empty(q.shape[:2] + (2, 2, 2))
    out[..., 0, 0, 0], out[..., 0, 0, 1] = a0, a3
    out[..., 0, 1, 0], out[..., 0, 1, 1] = a2, a1
    out[..., 1, 0, 0], out[..., 1, 0, 1] = -a2, a1
    out[..., 1, 1, 0], out[..., 1, 1, 1] = a0, -a3
    return out


# ---------------------------------------------------------------- device flow
class Flow:
    def __init__(self, dims=(8, 8, 8, 4), layers=8, bins=4, max_batch=64):
        d = _abi.FlowDesc()
        for k in range(4):
            d.n[k] = int(dims[k])
        d.layers, d.hidden, d.bins, d.max_batch = int(layers), HIDDEN, int(bins), int(max_batch)
        self.dims, self.layers, self.bins, self.max_batch = tuple(dims), layers, bins, max_batch
        self.V = int(np.prod(dims))
        self._h = ctypes.c_void_p()
        check(lib().qcd_flow_create(ctypes.byref(d), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(lib().qcd_flow_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def param_count(self):
        n = ctypes.c_size_t()
        check(lib().qcd_flow_param_count(self._h, ctypes.byref(n)))
        return n.value

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().qcd_flow_set_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size))

    def sample(self, seed, batch, beta):
        """(links (batch, 4, V, 4) float32, log q, action)."""
        links = np.empty((batch, 4, self.V, 4), dtype=np.float32)
        lq, S = np.empty(batch), np.empty(batch)
        D = ctypes.POINTER(ctypes.c_double)
        check(lib().qcd_flow_sample(self._h, int(seed), int(batch), float(beta),
                                    links.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lq.ctypes.data_as(D),
                                    S.ctypes.data_as(D)))
        return links, lq, S

    def train_grad(self, links):
        """Layer-local forward-KL loss per configuration and its gradient
        (summed over the batch) by the device's training kernels."""
        a = np.ascontiguousarray(links, dtype=np.float32).reshape(-1, 4, self.V, 4)
        loss = ctypes.c_double()
        g = np.empty(self.param_count(), dtype=np.float32)
        F = ctypes.POINTER(ctypes.c_float)
        check(lib().qcd_flow_train_grad(self._h, a.ctypes.data_as(F), a.shape[0], ctypes.byref(loss),
                                        g.ctypes.data_as(F)))
        return loss.value, g

    def train_step(self, links, lr):
        """One Adam step of the layer-local forward KL on `links`; returns
        the loss per configuration before the step."""
        a = np.ascontiguousarray(links, dtype=np.float32).reshape(-1, 4, self.V, 4)
        loss = ctypes.c_double()
        check(lib().qcd_flow_train_step(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.shape[0],
                                        float(lr), ctypes.byref(loss)))
        return loss.value

    def get_params(self):
        p = np.empty(self.param_count(), dtype=np.float32)
        check(lib().qcd_flow_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))
        return p

    def imh_sweep(self, lat, seed, beta):
        """One learned Metropolis sweep of `lat` (SU(2) FP32, the flow's
        extents) on the device: every coupling layer proposes its active
        links from its map of Haar draws and each is accepted with the
        exact Wilson-action Metropolis probability."""
        check(lib().qcd_flow_imh_sweep(self._h, lat._h, int(seed), float(beta)))

    def imh_sweep_next(self, lat, beta):
        """imh_sweep with the previous sweep's seed + 1 (the seed lives on
        the device, so a graph recorded around this call replays sweeps
        with successive seeds)."""
        check(lib().qcd_flow_imh_sweep_next(self._h, lat._h, float(beta)))

    def imh_seed(self, seed):
        """Sets the learned sweeps' seed without a sweep."""
        check(lib().qcd_flow_imh_seed(self._h, int(seed)))

    def imh_counts(self, reset=False):
        """(accepted, proposed) links of imh_sweep so far."""
        a, p = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().qcd_flow_imh_counts(self._h, ctypes.byref(a), ctypes.byref(p), 1 if reset else 0))
        return a.value, p.value

    def evaluate(self, links, beta):
        a = np.ascontiguousarray(links, dtype=np.float32).reshape(-1, 4, self.V, 4)
        lq, S = np.empty(a.shape[0]), np.empty(a.shape[0])
        D = ctypes.POINTER(ctypes.c_double)
        check(lib().qcd_flow_evaluate(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.shape[0],
                                      float(beta), lq.ctypes.data_as(D), S.ctypes.data_as(D)))
        return lq, S


# ---------------------------------------------------------------- training model
def _torch():
    import torch

    return torch


def geometry(dims):
    """Index tables: shift[d][+1/-1] (V,), stencil (V, 81), site coordinates (V, 4)."""
    n1, n2, n3, n4 = dims
    s = np.arange(n1 * n2 * n3 * n4)
    y = s % n2
    r = s // n2
    z = r % n3
    r //= n3
    t = r % n4
    x = r // n4
    c = np.stack([x, y, z, t], axis=1)
    n = np.array(dims)

    def idx(cc):
        cc = cc % n
        return cc[:, 1] + n2 * (cc[:, 2] + n3 * (cc[:, 3] + n4 * cc[:, 0]))

    shift = {}
    for d in range(4):
        for o in (1, -1):
            e = c.copy()
            e[:, d] += o
            shift[d, o] = idx(e)
    st = np.empty((len(s), STENCIL), dtype=np.int64)
    k = 0
    for ox in range(3):
        for oy in range(3):
            for oz in range(3):
                for ot in range(3):
                    st[:, k] = idx(c + np.array([ox - 1, oy - 1, oz - 1, ot - 1]))
                    k += 1
    return shift, st, c


def active_mask(coords, mu, par):
    s = par + sum(coords[:, d] for d in range(4) if d != mu)
    return (s % 2) == 0


class FlowModel:
    """FP32 PyTorch twin of the device flow (training and reference)."""

    def __init__(self, dims=(8, 8, 8, 4), layers=8, bins=4, device="cpu", seed=0):
        torch = _torch()
        self.dims, self.layers, self.bins, self.device = tuple(dims), layers, bins, device
        self.V = int(np.prod(dims))
        shift, st, coords = geometry(dims)
        T = lambda a: torch.as_tensor(a, device=device)  # noqa: E731
        self.shift = {k: T(v) for k, v in shift.items()}
        self.stencil = T(st)
        self.masks = [T(active_mask(coords, l % 4, (l // 4) % 2)) for l in range(layers)]
        g = torch.Generator().manual_seed(seed)
        p = torch.zeros(layers * LAYER_PARAMS)
        for l in range(layers):
            o = l * LAYER_PARAMS
            p[o:o + W1N] = torch.randn(W1N, generator=g) / math.sqrt(STENCIL * 5)
        self.params = p.to(device).requires_grad_(True)

    # views into the flat parameter vector (device layout)
    def layer_params(self, l):
        o = l * LAYER_PARAMS
        p = self.params
        W1 = p[o:o + W1N].view(HIDDEN, STENCIL * CIN0)
        b1 = p[o + W1N:o + W1N + HIDDEN]
        W2 = p[o + W1N + HIDDEN:o + W1N + HIDDEN + W2N].view(COUT, STENCIL * HIDDEN)
        b2 = p[o + W1N + HIDDEN + W2N:o + LAYER_PARAMS]
        return W1, b1, W2, b2

    def flat_params(self):
        return self.params.detach().cpu().numpy().astype(np.float32)

    # quaternion algebra on (..., 4)
    @staticmethod
    def qmul(a, b):
        torch = _torch()
        a0, a1, a2, a3 = a.unbind(-1)
        b0, b1, b2, b3 = b.unbind(-1)
        return torch.stack([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                            a0 * b1 + b0 * a1 - (a2 * b3 - a3 * b2),
                            a0 * b2 + b0 * a2 - (a3 * b1 - a1 * b3),
                            a0 * b3 + b0 * a3 - (a1 * b2 - a2 * b1)], dim=-1)

    @staticmethod
    def qdag(a):
        return a * a.new_tensor([1.0, -1.0, -1.0, -1.0])

    def staple(self, L, mu):
        """L: (B, 4, V, 4).  Staple sum of U_mu(x) at every site."""
        A = 0
        qm, qd, sh = self.qmul, self.qdag, self.shift
        for nu in range(4):
            if nu == mu:
                continue
            f = qm(qm(L[:, nu][:, sh[mu, 1]], qd(L[:, mu][:, sh[nu, 1]])), qd(L[:, nu]))
            xmn = sh[nu, -1]
            r = qm(qm(qd(L[:, nu][:, sh[mu, 1]][:, xmn]), qd(L[:, mu][:, xmn])), L[:, nu][:, xmn])
            A = A + f + r
        return A

    def plaq0(self, L, nu, rho):
        qm, qd, sh = self.qmul, self.qdag, self.shift
        p = qm(qm(L[:, nu], L[:, rho][:, sh[nu, 1]]), qm(qd(L[:, nu][:, sh[rho, 1]]), qd(L[:, rho])))
        return p[..., 0]

    def features(self, L, l):
        torch = _torch()
        mu = l % 4
        m = self.masks[l].to(L.dtype)
        A = self.staple(L, mu)
        k = torch.sqrt((A * A).sum(-1)) / 6.0 * m
        pl = [self.plaq0(L, nu, rho) for nu in range(4) for rho in range(nu + 1, 4) if mu not in (nu, rho)]
        z = torch.zeros_like(k)
        return torch.stack([k] + pl + [m.expand_as(k), z, z, z], dim=-1), A

    def conv(self, x, W, b):
        B, V, C = x.shape
        g = x[:, self.stencil]            # (B, V, 81, C)
        return g.reshape(B, V, STENCIL * C) @ W.t() + b

    def spline_params(self, r):
        torch = _torch()
        K = self.bins
        w = math.pi * (MIN_W + (1 - K * MIN_W) * torch.softmax(r[..., :K], -1))
        h = math.pi * (MIN_W + (1 - K * MIN_W) * torch.softmax(r[..., K:2 * K], -1))
        d = MIN_D + (1 - MIN_D) * torch.nn.functional.softplus(r[..., 2 * K:3 * K + 1]) / math.log(2.0)
        xk = torch.cat([torch.zeros_like(w[..., :1]), torch.cumsum(w, -1)], -1)
        yk = torch.cat([torch.zeros_like(h[..., :1]), torch.cumsum(h, -1)], -1)
        xk = torch.cat([xk[..., :-1], torch.full_like(xk[..., -1:], math.pi)], -1)
        yk = torch.cat([yk[..., :-1], torch.full_like(yk[..., -1:], math.pi)], -1)
        return xk, yk, d

    def spline(self, r, v, inverse):
        """Monotone rational-quadratic map on [0, pi]; returns (image, log f'(x))."""
        torch = _torch()
        xk, yk, d = self.spline_params(r)
        K = self.bins
        knots = yk if inverse else xk
        k = (v.unsqueeze(-1) >= knots[..., 1:K]).sum(-1).clamp(max=K - 1)
        gi = lambda a, j: torch.gather(a, -1, j.unsqueeze(-1)).squeeze(-1)  # noqa: E731
        x0, x1, y0, y1 = gi(xk, k), gi(xk, k + 1), gi(yk, k), gi(yk, k + 1)
        d0, d1 = gi(d, k), gi(d, k + 1)
        w, h = x1 - x0, y1 - y0
        s = h / w
        if not inverse:
            xi = ((v - x0) / w).clamp(0.0, 1.0)
        else:
            dy = v - y0
            dd = d1 + d0 - 2 * s
            a = h * (s - d0) + dy * dd
            bq = h * d0 - dy * dd
            cq = -s * dy
            disc = (bq * bq - 4 * a * cq).clamp(min=0.0)
            xi = (2 * cq / (-bq - torch.sqrt(disc))).clamp(0.0, 1.0)
        om = 1 - xi
        den = s + (d1 + d0 - 2 * s) * xi * om
        num = s * s * (d1 * xi * xi + 2 * s * xi * om + d0 * om * om)
        dlog = torch.log(num) - 2 * torch.log(den)
        out = x0 + xi * w if inverse else y0 + h * (s * xi * xi + d0 * xi * om) / den
        return out, dlog

    def layer(self, L, l, inverse):
        """One coupling layer on L (B, 4, V, 4); returns (L', log J per config)."""
        torch = _torch()
        mu = l % 4
        W1, b1, W2, b2 = self.layer_params(l)
        feat, A = self.features(L, l)
        hid = torch.tanh(self.conv(feat, W1, b1))
        r = self.conv(hid, W2, b2)
        m = self.masks[l]
        k = torch.sqrt((A * A).sum(-1)).clamp(min=1e-30)
        Vq = A / k.unsqueeze(-1)
        U = L[:, mu]
        W = self.qmul(U, Vq)
        sn = torch.sqrt((W[..., 1:] ** 2).sum(-1))
        th = torch.atan2(sn, W[..., 0])
        th2, dlog = self.spline(r, th, inverse)
        th_in, th_out = (th2, th) if inverse else (th, th2)
        si, so = torch.sin(th_in), torch.sin(th_out)
        good = (si > 1e-6) & (so > 1e-6)
        ratio_log = torch.where(good, torch.log(so.clamp(min=1e-30)) - torch.log(si.clamp(min=1e-30)), dlog)
        lj = torch.where(m, dlog + 2 * ratio_log, torch.zeros_like(dlog)).sum(-1)
        f = torch.sin(th2) / sn.clamp(min=1e-30)
        Wn = torch.cat([torch.cos(th2).unsqueeze(-1), W[..., 1:] * f.unsqueeze(-1)], -1)
        Un = self.qmul(Wn, self.qdag(Vq))
        Un = torch.where(m.unsqueeze(-1), Un, U)
        L = torch.cat([L[:, :mu], Un.unsqueeze(1), L[:, mu + 1:]], dim=1)
        return L, lj

    def local_nll(self, U):
        """The layer-local forward-KL loss per configuration: for every layer
        l, the sum over its active links of -log q_a(U_a | frozen links) with
        the layer applied to U itself (what the device's training kernels
        minimise)."""
        tot = 0
        for l in range(self.layers):
            _, lj = self.layer(U, l, True)
            tot = tot + lj
        return tot

    def forward(self, z):
        """Prior configurations -> (flow configurations, log q)."""
        logj = 0
        for l in range(self.layers):
            z, lj = self.layer(z, l, False)
            logj = logj + lj
        return z, -logj

    def log_q(self, U):
        """log q of given configurations (inverse flow)."""
        logj = 0
        for l in reversed(range(self.layers)):
            U, lj = self.layer(U, l, True)
            logj = logj + lj
        return -logj

    def action(self, L, beta):
        s = 0
        for nu in range(4):
            for rho in range(nu + 1, 4):
                s = s + (1.0 - self.plaq0(L, nu, rho)).sum(-1)
        return beta * s


def train_forward_kl(model, data, steps=1000, batch=32, lr=1e-3, seed=0, log_every=0):
    """Minimise -E_data[log q] (forward KL up to a constant) with Adam.
    data: (n, 4, V, 4) float32 array or tensor of heatbath configurations."""
    torch = _torch()
    dev = model.device
    X = torch.as_tensor(data, device=dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    opt = torch.optim.Adam([model.params], lr=lr)
    hist = []
    for it in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g).to(dev)
        loss = -model.log_q(X[idx]).mean()
        opt.zero_grad()
        loss.backward()
        torch.nn.utils.clip_grad_norm_([model.params], 10.0)
        opt.step()
        hist.append(float(loss.detach()))
        if log_every and (it % log_every == 0 or it == steps - 1):
            print("  step %5d  -<log q> %.3f" % (it, hist[-1]), flush=True)
    return hist


# ---------------------------------------------------------------- chains
def imh_chain(lat, flow, beta, steps, hb_per_step=1, batch=None, seed=1, rng=None, measure=True, wilson=True,
              log_weights=None):
    """Heatbath chain on `lat` (an SU(2) FP32 Lattice) with one
    independence-Metropolis step using `flow` (or none if flow is None)
    after every hb_per_step heatbath sweeps.  Returns (acceptance rate,
    seconds).  log_weights (a dict) collects log w = -S - log q of the
    current states ("current") and of the proposals ("proposal"): the
    acceptance is min(1, w(proposal) / w(current))."""
    rng = rng or np.random.default_rng(seed)
    batch = batch or (flow.max_batch if flow is not None else 1)
    pool, drawn, acc = [], 0, 0
    dims = lat.dims
    t0 = time.perf_counter()
    for step in range(steps):
        for _ in range(hb_per_step):
            lat.sweep("hb")
        if flow is not None:
            cur = to_quaternions(lat.export_links(), dims)
            lq_c, S_c = flow.evaluate(cur[None], beta)
            if not pool:
                links, _, S = flow.sample(seed * 1000003 + drawn, batch, beta)
                # log q of each proposal from the inverse pass, the same
                # density as the current state's, so that the acceptance is
                # one function of the two states
                lq, _ = flow.evaluate(links, beta)
                drawn += 1
                pool = [(links[b], lq[b], S[b]) for b in range(batch)]
            prop, lq_p, S_p = pool.pop()
            log_a = (-S_p + lq_c[0]) - (-S_c[0] + lq_p)
            if log_weights is not None:
                log_weights.setdefault("current", []).append(-S_c[0] - lq_c[0])
                log_weights.setdefault("proposal", []).append(-S_p - lq_p)
            if np.log(rng.random()) < log_a:
                lat.import_links(to_canonical(prop))
                acc += 1
        if measure:
            lat.measure(wilson=wilson)
    lat.sync()
    return acc / steps if flow is not None else 0.0, time.perf_counter() - t0
