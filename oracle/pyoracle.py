"""ctypes front end of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker (or as the timed CPU
baseline), never on the product path.

Besides the wrappers, three restatements are written here in numpy, each
pinned by its own tests: the reference's hot start (di_pierro, random_start,
produce_quads; tests/test_oracle_random_start.py), the twisted boundary
(twist_matrix, twisted_view, halfstep_twisted; tests/test_oracle_tbc.py) and
the per-slice action sums (action_slices; tests/test_oracle_action_slices.py).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

_F32P = ctypes.POINTER(ctypes.c_float)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)


class Lattice(ctypes.Structure):
    _fields_ = [
        ("sun", ctypes.c_int),
        ("n", ctypes.c_int * 4),
        ("beta", ctypes.c_double),
        ("nhit", ctypes.c_int),
    ]


class Ranlux(ctypes.Structure):
    _fields_ = [
        ("s", ctypes.c_float * 24),
        ("i24", ctypes.c_int),
        ("j24", ctypes.c_int),
        ("in24", ctypes.c_int),
        ("carry", ctypes.c_float),
        ("nskip", ctypes.c_int),
    ]


class Rng(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("seq_generator", ctypes.c_int),
        ("ranlux", ctypes.POINTER(Ranlux)),
        ("pm_seed", _I32P),
        ("quads", _F32P),
        ("stride", ctypes.c_ulong),
    ]


RNG_GID, RNG_SEQ, RNG_QUADS = 0, 1, 2
HB, OR = 0, 1


def build():
    """Compile liboracle.so with gcc (the oracle is plain C)."""
    import subprocess

    subprocess.run(
        ["gcc", "-O2", "-shared", "-fPIC", os.path.join(HERE, "oracle.c"), "-o", LIB_PATH, "-lm"],
        check=True,
    )


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_xor128_step.restype = ctypes.c_uint32
        L.orc_pm_step.restype = ctypes.c_uint32
        L.orc_pm_step.argtypes = [ctypes.c_uint32]
        L.orc_pm_one_cpu.restype = ctypes.c_float
        L.orc_ranlux_next.restype = ctypes.c_float
        L.orc_sites.restype = ctypes.c_uint
        L.orc_neighbour.restype = ctypes.c_uint
        L.orc_parity_site.restype = ctypes.c_uint
        L.orc_link_action.restype = ctypes.c_double
        L.orc_wilson.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def lattice(sun, dims, beta=5.9, nhit=10):
    L = Lattice()
    L.sun = sun
    for i in range(4):
        L.n[i] = dims[i]
    L.beta = beta
    L.nhit = nhit
    return L


def link_shape(L):
    V = int(np.prod(list(L.n)))
    return (V, 4, L.sun, L.sun, 2)


def cold(L):
    a = np.zeros(link_shape(L))
    lib().orc_init_cold(ctypes.byref(L), _p(a, _F64P))
    return a


def gid(L):
    a = np.zeros(link_shape(L))
    lib().orc_init_gid(ctypes.byref(L), _p(a, _F64P))
    return a


def hot(L, quads, stride):
    """Haar start from per-stream quads (quads[k, i, :] = quad k of stream i,
    as returned by the produce functions with n = stride instances)."""
    q = np.ascontiguousarray(quads, dtype=np.float32)
    a = np.zeros(link_shape(L))
    lib().orc_init_hot(ctypes.byref(L), _p(a, _F64P), _p(q, _F32P), stride)
    return a


def di_pierro(u, v, w):
    """lattice_random2 (su2_update_cl.cl:110-146): M. Di Pierro's map of
    three uniform numbers to SU(2), alpha = pi u, phi = 2 pi v,
    cos theta = 2 w - 1, quaternion a = (cos alpha, sin alpha sin theta
    cos phi, sin alpha sin theta sin phi, sin alpha cos theta) as
    [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]].  Arrays in, (..., 2, 2)
    complex out, in float64."""
    u, v, w = (np.asarray(x, dtype=np.float64) for x in (u, v, w))
    alpha, phi, c = np.pi * u, 2 * np.pi * v, 2 * w - 1
    t = np.sin(alpha) * np.sqrt(1 - c * c)
    a0, a1, a2, a3 = np.cos(alpha), t * np.cos(phi), t * np.sin(phi), np.sin(alpha) * c
    m = np.empty(u.shape + (2, 2), dtype=np.complex128)
    m[..., 0, 0] = a0 + 1j * a3
    m[..., 0, 1] = a2 + 1j * a1
    m[..., 1, 0] = -a2 + 1j * a1
    m[..., 1, 1] = a0 - 1j * a3
    return m


def random_start(L, bufs, rowsize):
    """The reference's hot start (lattice_init_hot_X..T, suncl.cl:56-161).
    bufs[d]: the quads of the produce() before direction d, flat ((samples *
    instances, 4), quad q of the reference's buffer at row q).  The link of
    site s (this oracle's site order, which is the reference's table index)
    takes quad s (SU(2)) or quads s, s + rowsize, s + 2 rowsize (SU(3)): the
    SU(2) elements A, B, C of components x, y, z (u from the first quad, v
    from the second, w from the third) multiplied as E12(A) E13(B) E23(C)
    (lattice_random3, su3_update_cl.cl:165-225)."""
    V = int(np.prod(list(L.n)))
    N = L.sun
    s = np.arange(V)
    U = np.zeros((V, 4, N, N), dtype=np.complex128)
    for d in range(4):
        b = np.asarray(bufs[d], dtype=np.float32).reshape(-1, 4)
        if N == 2:
            q = b[s]
            U[:, d] = di_pierro(q[:, 0], q[:, 1], q[:, 2])
            continue
        q0, q1, q2 = b[s], b[s + rowsize], b[s + 2 * rowsize]
        m = np.tile(np.eye(3, dtype=np.complex128), (V, 1, 1))
        for c, (i, j) in enumerate(((0, 1), (0, 2), (1, 2))):
            w = di_pierro(q0[:, c], q1[:, c], q2[:, c])
            e = np.tile(np.eye(3, dtype=np.complex128), (V, 1, 1))
            e[:, i, i], e[:, i, j], e[:, j, i], e[:, j, j] = w[:, 0, 0], w[:, 0, 1], w[:, 1, 0], w[:, 1, 1]
            m = m @ e
        U[:, d] = m
    return np.ascontiguousarray(np.stack([U.real, U.imag], axis=-1))


def produce_quads(L, hot_run, instances=None):
    """Quads per instance of one produce() (suncl.cpp:2302-2307,
    random.cpp:219-223): align64 of 3 PRNGSTEP (NHIT + 1) quads, a hot run
    adding ROWSIZE * ND, ROWSIZE = align64(V), over PRNG_instances
    (default PRNGSTEP = align64(V/2)): PRNG_samples = 1 + (total - 1) /
    instances."""
    V = int(np.prod(list(L.n)))
    prngstep = -(-(V // 2) // 64) * 64
    rowsize = -(-V // 64) * 64
    total = 3 * prngstep * (L.nhit + 1) + (4 * rowsize if hot_run else 0)
    total = -(-total // 64) * 64
    return 1 + (total - 1) // (instances or prngstep)


def rand_fill(seed, count):
    out = np.zeros(count, dtype=np.int32)
    lib().orc_rand_fill(seed, count, _p(out, _I32P))
    return out


def xor128_seed_table(randseries, n):
    t = np.zeros((n, 4), dtype=np.uint32)
    lib().orc_xor128_seed_table(randseries, n, _p(t, _U32P))
    return t


def xor128_produce(table, samples):
    n = table.shape[0]
    q = np.zeros((samples, n, 4), dtype=np.float32)
    lib().orc_xor128_produce(_p(table, _U32P), n, samples, _p(q, _F32P))
    return q


def pm_seed_table(randseries, n):
    t = np.zeros((n, 4), dtype=np.uint32)
    lib().orc_pm_seed_table(randseries, n, _p(t, _U32P))
    return t


def pm_produce(table, samples):
    n = table.shape[0]
    q = np.zeros((samples, n, 4), dtype=np.float32)
    lib().orc_pm_produce(_p(table, _U32P), n, samples, _p(q, _F32P))
    return q


def pm_cpu_stream(seed, count):
    s = ctypes.c_int32(seed)
    return np.array([lib().orc_pm_one_cpu(ctypes.byref(s)) for _ in range(count)], dtype=np.float32)


def ranlux_nskip(lux):
    return lib().orc_ranlux_nskip(lux)


def ranlux_cpu_stream(randseries, nskip, count):
    out = np.zeros(count, dtype=np.float32)
    lib().orc_ranlux_cpu_stream(randseries, nskip, count, _p(out, _F32P))
    return out


def ranlux_seed_table(randseries, n):
    t = np.zeros((7, n, 4), dtype=np.float32)
    lib().orc_ranlux_seed_table(randseries, n, _p(t, _F32P))
    return t


def ranlux_produce(table, samples, nskip):
    n = table.shape[1]
    q = np.zeros((samples, n, 4), dtype=np.float32)
    lib().orc_ranlux_produce(_p(table, _F32P), n, samples, nskip, _p(q, _F32P))
    return q


class SeqRng:
    """Sequential scalar source (suncpp semantics) after srand(randseries)."""

    def __init__(self, generator, randseries, nskip=223):
        self.rng = Rng()
        self.rng.kind = RNG_SEQ
        if generator == "PM":
            lib().orc_srand(randseries)
            self._seed = ctypes.c_int32(lib().orc_rand() % 2147483647)
            self.rng.seq_generator = 1
            self.rng.pm_seed = ctypes.pointer(self._seed)
        else:
            self._rl = Ranlux()
            lib().orc_srand(randseries)
            jseed = lib().orc_rand() % 2147483647
            lib().orc_ranlux_init_scalar(ctypes.byref(self._rl), jseed, nskip)
            self.rng.seq_generator = 0
            self.rng.ranlux = ctypes.pointer(self._rl)


class QuadRng:
    """Reference GPU buffer semantics: thread i reads quads[i + k*stride]."""

    def __init__(self, quads, stride):
        self._q = np.ascontiguousarray(quads, dtype=np.float32)
        self.rng = Rng()
        self.rng.kind = RNG_QUADS
        self.rng.quads = _p(self._q, _F32P)
        self.rng.stride = stride


class GidRng:
    def __init__(self):
        self.rng = Rng()
        self.rng.kind = RNG_GID


def halfstep(L, links, parity, mu, mode, rng):
    return lib().orc_halfstep(ctypes.byref(L), _p(links, _F64P), parity, mu, mode, ctypes.byref(rng.rng))


def halfstep_src(L, links, src, parity, mu, mode, rng):
    """orc_halfstep with the staples read from `src` (parity 2: both
    parities in one pass)."""
    return lib().orc_halfstep_src(ctypes.byref(L), _p(links, _F64P), _p(src, _F64P), parity, mu, mode,
                                  ctypes.byref(rng.rng))


def sweep(L, links, mode, rng):
    return lib().orc_sweep(ctypes.byref(L), _p(links, _F64P), mode, ctypes.byref(rng.rng))


def wilson(L, links, R, T):
    """Sum of ReTr W(R x T) over sites and the three spatial directions."""
    return lib().orc_wilson(ctypes.byref(L), _p(np.ascontiguousarray(links), _F64P), R, T)


def plaquette(L, links):
    out = np.zeros(2)
    lib().orc_plaquette(ctypes.byref(L), _p(links, _F64P), _p(out, _F64P))
    return out


def fmunu(L, links, efield=False, gen1=3, gen2=8):
    """F_munu sums as complex [generator][plane]."""
    out = np.zeros(12)
    lib().orc_fmunu(ctypes.byref(L), _p(links, _F64P), int(efield), int(gen1), int(gen2), _p(out, _F64P))
    return (out[0::2] + 1j * out[1::2]).reshape(2, 3)


def plaquette_sites(L, links):
    """(V, 2): per-site sums of the spatial / temporal ReTr U_p."""
    V = int(np.prod(list(L.n)))
    out = np.zeros((V, 2))
    lib().orc_plaquette_sites(ctypes.byref(L), _p(np.ascontiguousarray(links), _F64P), _p(out, _F64P))
    return out


def action_slices(L, links):
    """Per x / y / z slice sums of plaquette_sites: dict -> (N_d, 2)."""
    n1, n2, n3, n4 = list(L.n)
    a = plaquette_sites(L, links).reshape(n1, n4, n3, n2, 2)   # site order: x slowest, then t, z, y
    return {"x": a.sum(axis=(1, 2, 3)), "y": a.sum(axis=(0, 1, 2)), "z": a.sum(axis=(0, 1, 3))}


def polyakov(L, links):
    out = np.zeros(4)
    lib().orc_polyakov(ctypes.byref(L), _p(links, _F64P), _p(out, _F64P))
    return out


def staple(L, links, site, mu):
    A = np.zeros((L.sun, L.sun, 2))
    lib().orc_staple(ctypes.byref(L), _p(links, _F64P), site, mu, _p(A, _F64P))
    return A


def link_action(L, links, site, mu):
    return lib().orc_link_action(ctypes.byref(L), _p(links, _F64P), site, mu)


def reunitarize(L, links):
    lib().orc_reunitarize(ctypes.byref(L), _p(links, _F64P))


def parity_site(L, i, parity):
    return lib().orc_parity_site(ctypes.byref(L), i, parity)


def neighbour(L, site, d, sign):
    return lib().orc_neighbour(ctypes.byref(L), site, d, sign)


# ---------------------------------------------------------------- twisted boundary
def twist_matrix(sun, phi, omega=0.0):
    """The matrix T that lattice_table_3 / lattice_table_2 under TBC
    (su3_matrix_memory.cl:57-98, su2_matrix_memory.cl:56-88) multiplies the
    Y link at x = N1 - 1 by, from the left: SU(3) scales the stored first row
    by e^{i(phi+omega)/2} and the second by e^{i(omega-phi)/2} (the rebuilt
    third row then carries e^{-i omega}); SU(2) multiplies by
    Omega = diag(e^{i phi/2}, e^{-i phi/2}) (omega unused)."""
    if sun == 3:
        d = np.exp(1j * np.array([(phi + omega) / 2, (omega - phi) / 2, -omega]))
    else:
        d = np.exp(1j * np.array([phi / 2, -phi / 2]))
    return np.diag(d)


def _x_of_sites(L):
    """x coordinate of every site in the oracle's order (x slowest)."""
    V = int(np.prod(list(L.n)))
    return np.arange(V) // (L.n[1] * L.n[2] * L.n[3])


def twisted_view(L, links, phi, omega=0.0):
    """The links as every read under TBC sees them: Y links (direction 1) of
    the sites with x = N1 - 1 replaced by T U; all others unchanged."""
    out = np.array(links, dtype=np.float64, copy=True)
    if phi == 0.0 and omega == 0.0:
        return out
    T = twist_matrix(L.sun, phi, omega)
    sel = _x_of_sites(L) == L.n[0] - 1
    U = out[sel, 1, ..., 0] + 1j * out[sel, 1, ..., 1]
    W = T @ U
    out[sel, 1, ..., 0], out[sel, 1, ..., 1] = W.real, W.imag
    return out


def halfstep_twisted(L, links, parity, mu, mode, rng, phi, omega=0.0):
    """One (parity, mu) half-step under TBC: every read sees the twisted view
    (the link being updated included), the results are stored as computed
    (update_{even,odd}_Y with lattice_table_3 / lattice_store_3,
    suncl.cl:505-552).  A half-step reads no link it writes (staples of
    (parity, mu) links hold only direction-mu links of the other parity), so
    running the plain half-step on the view and copying its (parity, mu)
    links back is that rule exactly."""
    view = twisted_view(L, links, phi, omega)
    rejected = halfstep(L, view, parity, mu, mode, rng)
    V = int(np.prod(list(L.n)))
    sites = np.array([parity_site(L, i, parity) for i in range(V // 2)])
    links[sites, mu] = view[sites, mu]
    return rejected
