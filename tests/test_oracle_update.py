"""Pins the oracle's update and measurement arithmetic.

Golden links come from the reference's own suncpp run on this machine
(tests/golden/make_golden.py): GID known-answer mode (INTS = 2) and the
sequential RANLUX/PM streams.  The remaining functions are pinned by what
the mathematics fixes: closed forms on cold starts, gauge invariance, gauge
covariance of the staple, the Kennedy-Pendleton density, and the
overrelaxation invariants (no reference code exists for overrelaxation).
"""
import os

import numpy as np
import pytest

# (tag, sun, dims, beta, nhit, ints, generator, randseries, kept sweeps)
CASES = [
    ("su3_gid_4444", 3, (4, 4, 4, 4), 5.9, 10, 2, "RANLUX3", 54, (1, 2)),
    ("su3_gid_6424", 3, (6, 4, 2, 4), 5.9, 10, 2, "RANLUX3", 54, (1,)),
    ("su3_cold_rl3_6424", 3, (6, 4, 2, 4), 5.9, 10, 1, "RANLUX3", 54, (1, 2)),
    ("su3_cold_pm_4444", 3, (4, 4, 4, 4), 6.0, 10, 1, "PM", 7, (1,)),
    ("su2_gid_4444", 2, (4, 4, 4, 4), 2.3, 10, 2, "RANLUX3", 54, (1, 2)),
    ("su2_cold_rl3_6424", 2, (6, 4, 2, 4), 2.3, 10, 1, "RANLUX3", 54, (1, 2)),
    ("su2_cold_rl0_4444", 2, (4, 4, 4, 4), 1.7, 2, 1, "RANLUX0", 1, (1,)),
]


def _rng(O, ints, gen, seed):
    if ints == 2:
        return O.GidRng()
    if gen == "PM":
        return O.SeqRng("PM", seed)
    return O.SeqRng("RANLUX", seed, O.ranlux_nskip(int(gen[-1])))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_sweeps_match_reference_suncpp(O, golden, case):
    tag, sun, dims, beta, nhit, ints, gen, seed, keep = case
    L = O.lattice(sun, dims, beta, nhit)
    a = O.gid(L) if ints == 2 else O.cold(L)
    rng = _rng(O, ints, gen, seed)
    plq_ref = np.loadtxt(os.path.join(golden, "sweep_%s_plq.txt" % tag))
    V = int(np.prod(dims))
    for k in range(1, max(keep) + 1):
        O.sweep(L, a, O.HB, rng)
        p = O.plaquette(L, a) / (3 * V)
        assert abs(p[0] - plq_ref[k, 1]) <= 1e-13 * abs(plq_ref[k, 1])
        assert abs(p[1] - plq_ref[k, 2]) <= 1e-13 * abs(plq_ref[k, 2])
        if k in keep:
            ref = np.load(os.path.join(golden, "sweep_%s_s%d.npy" % (tag, k)))
            assert np.abs(a - ref).max() <= 1e-13


def test_gid_start_matches_reference(O, golden):
    for tag, sun, dims in (("su3_gid_4444", 3, (4, 4, 4, 4)), ("su2_gid_4444", 2, (4, 4, 4, 4))):
        L = O.lattice(sun, dims)
        plq = np.loadtxt(os.path.join(golden, "sweep_%s_plq.txt" % tag))
        p = O.plaquette(L, O.gid(L)) / (3 * np.prod(dims))
        assert np.allclose(p, plq[0, 1:3], rtol=1e-14)


def _random_su(rng, N, shape):
    z = rng.normal(size=shape + (N, N)) + 1j * rng.normal(size=shape + (N, N))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / N))[..., None, None]
    return q


def _to_links(U):
    return np.stack([U.real, U.imag], axis=-1).copy()


def _from_links(a):
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("sun", [2, 3])
def test_cold_start_closed_forms(O, sun):
    L = O.lattice(sun, (4, 4, 2, 6))
    a = O.cold(L)
    V = 4 * 4 * 2 * 6
    assert np.allclose(O.plaquette(L, a), [3 * V * sun, 3 * V * sun])
    poly = O.polyakov(L, a)
    Vs = V // 6
    assert np.allclose(poly, [Vs * sun, 0.0, Vs * sun ** 2, Vs * sun ** 4])
    # staple of a cold configuration is 6 * identity
    A = _from_links(O.staple(L, a, 5, 2))
    assert np.allclose(A, 6 * np.eye(sun))


def _gauge_transform(U, g, L):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dagger."""
    V = U.shape[0]
    out = np.empty_like(U)
    for s in range(V):
        for mu in range(4):
            nb = _neighbour(L, s, mu)
            out[s, mu] = g[s] @ U[s, mu] @ g[nb].conj().T
    return out


def _neighbour(L, s, mu):
    import pyoracle

    return pyoracle.neighbour(L, s, mu, 1)


@pytest.mark.parametrize("sun", [2, 3])
def test_measurements_gauge_invariant(O, sun):
    rng = np.random.default_rng(11)
    dims = (4, 2, 4, 6)
    L = O.lattice(sun, dims)
    V = int(np.prod(dims))
    U = _random_su(rng, sun, (V, 4))
    g = _random_su(rng, sun, (V,))
    a = _to_links(U)
    b = _to_links(_gauge_transform(U, g, L))
    assert np.allclose(O.plaquette(L, a), O.plaquette(L, b), rtol=1e-12)
    assert np.allclose(O.polyakov(L, a), O.polyakov(L, b), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("sun", [2, 3])
def test_plaquette_brute_force(O, sun):
    """Independent numpy evaluation of the six plane sums."""
    rng = np.random.default_rng(5)
    dims = (2, 4, 2, 4)
    L = O.lattice(sun, dims)
    V = int(np.prod(dims))
    U = _random_su(rng, sun, (V, 4))
    sp = tm = 0.0
    for s in range(V):
        for a_, b_ in ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3)):
            xa, xb = _neighbour(L, s, a_), _neighbour(L, s, b_)
            t = np.trace(U[s, a_] @ U[xa, b_] @ U[xb, a_].conj().T @ U[s, b_].conj().T).real
            if b_ == 3:
                tm += t
            else:
                sp += t
    assert np.allclose(O.plaquette(L, _to_links(U)), [sp, tm], rtol=1e-12)


def test_polyakov_brute_force(O):
    rng = np.random.default_rng(8)
    dims = (2, 2, 4, 4)
    L = O.lattice(3, dims)
    V = int(np.prod(dims))
    U = _random_su(rng, 3, (V, 4))
    re = im = p2 = p4 = 0.0
    import pyoracle

    for x in range(dims[0]):
        for y in range(dims[1]):
            for z in range(dims[2]):
                P = np.eye(3, dtype=complex)
                for t in range(dims[3]):
                    s = y + dims[1] * (z + dims[2] * (t + dims[3] * x))
                    P = P @ U[s, 3]
                tr = np.trace(P)
                re += tr.real
                im += tr.imag
                p2 += abs(tr) ** 2
                p4 += abs(tr) ** 4
    assert np.allclose(O.polyakov(L, _to_links(U)), [re, im, p2, p4], rtol=1e-12)


@pytest.mark.parametrize("sun", [2, 3])
def test_staple_gauge_covariance(O, sun):
    """A_mu(x) -> g(x+mu) A_mu(x) g(x)^dagger, so ReTr(U A) is invariant."""
    rng = np.random.default_rng(3)
    dims = (4, 4, 2, 2)
    L = O.lattice(sun, dims)
    V = int(np.prod(dims))
    U = _random_su(rng, sun, (V, 4))
    g = _random_su(rng, sun, (V,))
    Ug = _gauge_transform(U, g, L)
    for s, mu in ((0, 0), (7, 1), (13, 2), (30, 3)):
        A = _from_links(O.staple(L, _to_links(U), s, mu))
        Ag = _from_links(O.staple(L, _to_links(Ug), s, mu))
        xm = _neighbour(L, s, mu)
        assert np.allclose(Ag, g[xm] @ A @ g[s].conj().T, atol=1e-12)
        assert np.isclose(O.link_action(L, _to_links(U), s, mu), O.link_action(L, _to_links(Ug), s, mu))


@pytest.mark.parametrize("sun", [2, 3])
def test_overrelaxation_invariants(O, sun):
    """OR keeps every updated link's ReTr(U A) (staple fixed within a
    half-step), keeps links unitary, and for SU(2) is an involution."""
    rng = np.random.default_rng(21)
    dims = (4, 4, 2, 4)
    L = O.lattice(sun, dims, beta=5.9 if sun == 3 else 2.3)
    V = int(np.prod(dims))
    a = _to_links(_random_su(rng, sun, (V, 4)))
    parity, mu = 1, 2
    sites = [O.parity_site(L, i, parity) for i in range(V // 2)]
    before = [O.link_action(L, a, s, mu) for s in sites]
    b = a.copy()
    O.halfstep(L, b, parity, mu, O.OR, O.GidRng())
    after = [O.link_action(L, b, s, mu) for s in sites]
    assert np.allclose(before, after, rtol=1e-12, atol=1e-12)
    Ub = _from_links(b)
    eye = np.eye(sun)
    assert np.abs(np.einsum("...ji,...jk->...ik", Ub.conj(), Ub) - eye).max() < 1e-13
    changed = np.abs(b - a).max()
    assert changed > 1e-3
    if sun == 2:
        c = b.copy()
        O.halfstep(L, c, parity, mu, O.OR, O.GidRng())
        assert np.abs(c - a).max() < 1e-12


def test_heatbath_unitarity_and_locality(O):
    """A half-step changes only links (parity, mu) and leaves them SU(3)."""
    L = O.lattice(3, (4, 4, 4, 4), 5.9)
    a = O.cold(L)
    b = a.copy()
    rng = O.SeqRng("RANLUX", 54, 223)
    O.halfstep(L, b, 0, 1, O.HB, rng)
    V = 256
    for s in range(V):
        for mu in range(4):
            c = [int(v) for v in np.unravel_index(s, (4, 4, 4, 4))]
            par = sum(c) & 1
            moved = np.abs(b[s, mu] - a[s, mu]).max() > 0
            assert moved == (mu == 1 and par == 0)
    U = _from_links(b)
    assert np.abs(np.einsum("...ji,...jk->...ik", U.conj(), U) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_kennedy_pendleton_distribution(O):
    """Single SU(2) link in a fixed staple A = k 1: the new link's a0 has
    density ~ sqrt(1-a0^2) exp(beta_eff k a0), mean I_2(x)/I_1(x) with
    x = beta_eff * 2k (KP draws exp(beta*det*a0), det = |quaternion M| = 2k)."""
    from scipy.special import iv

    L = O.lattice(2, (2, 2, 2, 2), beta=2.0, nhit=50)
    a = O.cold(L)   # staple of every link = 6 * identity
    rng = O.SeqRng("RANLUX", 54, 223)
    vals = []
    for _ in range(400):
        b = a.copy()
        O.halfstep(L, b, 0, 0, O.HB, rng)
        for i in range(8):
            s = O.parity_site(L, i, 0)
            vals.append(b[s, 0, 0, 0, 0])
    vals = np.array(vals)
    x = (2.0 / 2) * 12.0
    mean = iv(2, x) / iv(1, x)
    assert abs(vals.mean() - mean) < 4 * vals.std() / np.sqrt(len(vals))


def _quads(stride, nq, fill, nan_from):
    q = np.empty((nq + 4, stride, 4), dtype=np.float32)
    q[:] = np.float32(np.nan)
    q[:nan_from] = np.array(fill, dtype=np.float32)
    return q


def test_quads_fetch_rule_accept(O):
    """GPU buffer semantics (su3_update_cl.cl:597-668): with every first try
    accepted, a link consumes exactly 2 quads per SU(2) subgroup (the try and
    the accept quad) at i + k*stride.  Quads beyond k = 6 are NaN, so a wider
    read would poison the link."""
    L = O.lattice(3, (2, 2, 2, 2), 5.9)
    a = O.cold(L)
    q = _quads(8, 6, (0.3, 0.2, 0.4, 0.0), 6)
    O.halfstep(L, a, 0, 0, O.HB, O.QuadRng(q, 8))
    assert np.isfinite(a).all()
    assert np.abs(a - O.cold(L)).max() > 1e-3


@pytest.mark.parametrize("nhit", [1, 3, 10])
def test_quads_fetch_rule_reject_skips_one(O, nhit):
    """All NHIT tries rejected: each subgroup consumes NHIT tries plus one
    skipped quad, so a link reads exactly 3*(NHIT+1) quads and is left
    unchanged (identity in every subgroup of a cold link)."""
    L = O.lattice(3, (2, 2, 2, 2), 5.9, nhit)
    a = O.cold(L)
    n = 3 * (nhit + 1)
    q = _quads(8, n, (0.999999, 0.0, 0.999999, 1.0), n)
    rej = O.halfstep(L, a, 1, 2, O.HB, O.QuadRng(q, 8))
    assert rej == 8 * 3
    assert np.isfinite(a).all()
    assert np.abs(a - O.cold(L)).max() < 1e-15
    # one quad fewer is not enough: the last skip of the last subgroup would
    # read the NaN-free region only if consumption were smaller
    q2 = _quads(8, n, (0.999999, 0.0, 0.999999, 1.0), n - 1)
    b = O.cold(L)
    O.halfstep(L, b, 1, 2, O.HB, O.QuadRng(q2, 8))
    assert np.isfinite(b).all()   # the skipped quad's values are never used


def test_quads_first_reject_then_accept(O):
    """Mixed: try 1 rejected, try 2 accepted -> 3 quads per subgroup."""
    L = O.lattice(3, (2, 2, 2, 2), 5.9, 10)
    a = O.cold(L)
    q = np.full((12, 8, 4), np.nan, dtype=np.float32)
    for sg in range(3):
        q[3 * sg + 0] = (0.999999, 0.0, 0.999999, 1.0)
        q[3 * sg + 1] = (0.3, 0.2, 0.4, 0.0)
        q[3 * sg + 2] = (0.6, 0.7, 0.0, 0.0)
    rej = O.halfstep(L, a, 0, 3, O.HB, O.QuadRng(q, 8))
    assert rej == 0
    assert np.isfinite(a).all()


# ---------------------------------------------------------------- Haar hot start
def _moment_ok(x, exact, nsig=5.0):
    se = np.std(x) / np.sqrt(len(x))
    return abs(np.mean(x) - exact) <= nsig * max(se, 1e-12), (np.mean(x), exact, se)


@pytest.mark.parametrize("sun", [3, 2])
def test_hot_start_is_haar(O, sun):
    """Haar moments of the oracle's hot start (independent of its formula):
    SU(N): E Tr U = 0, E |Tr U|^2 = 1, E |Tr U|^4 = 2 (Diaconis-Shahshahani;
    for SU(2) Tr U = 2 a0 with a0 semicircular), and for SU(3)
    E (Tr U)^3 = 1 (the baryon-like invariant that vanishes for U(3))."""
    dims = (12, 12, 12, 12)
    L = O.lattice(sun, dims, 5.9, 10)
    n = int(np.prod(dims)) // 2
    Q = 3 if sun == 3 else 1
    q = O.xor128_produce(O.xor128_seed_table(7, n), 8 * Q)
    a = O.hot(L, q, n)
    M = (a[..., 0] + 1j * a[..., 1]).reshape(-1, sun, sun)
    uu = np.einsum("nij,nkj->nik", M, M.conj())
    assert np.abs(uu - np.eye(sun)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13
    tr = np.trace(M, axis1=1, axis2=2)
    for x, exact in [(tr.real, 0.0), (tr.imag, 0.0), (np.abs(tr) ** 2, 1.0), (np.abs(tr) ** 4, 2.0)]:
        ok, info = _moment_ok(x, exact)
        assert ok, info
    if sun == 3:
        ok, info = _moment_ok((tr ** 3).real, 1.0)
        assert ok, info
        ok, info = _moment_ok((tr ** 3).imag, 0.0)
        assert ok, info


def test_hot_start_box_muller_edge(O):
    """u1 = 1 (XOR128 output 2^32 after rounding) must not give inf/nan."""
    L = O.lattice(2, (2, 2, 2, 2), 2.3, 10)
    n = 8
    q = np.ones((8, n, 4), dtype=np.float32)
    a = O.hot(L, q, n)
    assert np.isfinite(a).all()


# ---------------------------------------------------------------- Wilson loop
@pytest.mark.parametrize("sun", [3, 2])
def test_wilson_loop_pins(O, sun):
    """W(1,1) summed over sites and spatial directions is the temporal
    plaquette sum (each term is a temporal plaquette traversed the other way,
    same real trace); a cold lattice gives 3 V N for any R, T; gauge
    invariance under random g(x); W(R,T) by explicit matrix products."""
    dims = (4, 4, 4, 4)
    L = O.lattice(sun, dims, 5.9, 10)
    a = O.gid(L)
    V = int(np.prod(dims))
    assert np.isclose(O.wilson(L, a, 1, 1), O.plaquette(L, a)[1], rtol=1e-12)
    c = O.cold(L)
    for R, T in [(1, 1), (2, 3), (3, 2)]:
        assert np.isclose(O.wilson(L, c, R, T), 3 * V * sun, rtol=1e-14)
    # gauge invariance
    rng = np.random.default_rng(3)
    M = (a[..., 0] + 1j * a[..., 1]).reshape(V, 4, sun, sun)
    g = np.linalg.qr(rng.normal(size=(V, sun, sun)) + 1j * rng.normal(size=(V, sun, sun)))[0]
    Mg = M.copy()
    for s in range(V):
        for d in range(4):
            Mg[s, d] = g[s] @ M[s, d] @ g[O.neighbour(L, s, d, +1)].conj().T
    ag = np.stack([Mg.real, Mg.imag], axis=-1).reshape(a.shape)
    assert np.isclose(O.wilson(L, ag, 2, 2), O.wilson(L, a, 2, 2), rtol=1e-11)
    # explicit loop at one site and direction, summed by brute force
    def U(site, d):
        return M[site, d]

    def step(site, d, k=1):
        for _ in range(k):
            site = O.neighbour(L, site, d, +1)
        return site

    R, T = 2, 2
    tot = 0.0
    for x in range(V):
        B = U(x, 3) @ U(step(x, 3), 3)
        for d in range(3):
            xt, xr = step(x, 3, T), step(x, d, R)
            Rm = U(xt, d) @ U(step(xt, d), d)
            Tm = U(xr, 3) @ U(step(xr, 3), 3)
            Lm = U(x, d) @ U(step(x, d), d)
            tot += np.trace(B @ Rm @ Tm.conj().T @ Lm.conj().T).real
    assert np.isclose(O.wilson(L, a, 2, 2), tot, rtol=1e-12)
