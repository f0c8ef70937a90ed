"""The .qcg state file: the oracle's restatement of the reference layout
(oracle/qcg.py, suncl.cpp:1857-2041) pinned by arithmetic and unitarity, and
the library's header reader checked against files the oracle writes.  No
device call: these run without a GPU."""
import numpy as np
import pytest

import qcg
from qcdgpu_amd import QCDError, _abi, read_qcg_header


def haar(rng, V, N):
    z = rng.normal(size=(V, 4, N, N)) + 1j * rng.normal(size=(V, 4, N, N))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(q)
    q[..., :, 0] /= det[..., None]   # det 1 without leaving U(N)
    return np.stack([q.real, q.imag], axis=-1)


def test_string_words():
    # "QCDG" = 0x51, 0x43, 0x44, 0x47 low byte first; "PU" = 0x50, 0x55; "1.5"
    assert qcg.string_word("QCDGPU", 0) == 0x47444351
    assert qcg.string_word("QCDGPU", 4) == 0x00005550
    assert qcg.string_word("1.5", 0) == 0x00352E31


def test_section_sizes():
    # 32^3 x 8, ITER 500: ROWSIZE = V, table 12 V float4, partials V / 64
    s = qcg.sizes(3, (32, 32, 32, 8), 500)
    assert s["rowsize"] == 262144 and s["table_size"] == 12 * 262144
    assert s["measurement_size_F"] == 4096 and s["energies_size"] == 512 == s["polyakov_loop_size"]
    # 6^3 x 3 = 648 sites: ROWSIZE 704, table 12 * 704 = 8448, partials
    # ceil(704 / 64) = 11 aligned to 64; SU(2) table 4 * 704
    s = qcg.sizes(3, (6, 6, 6, 3), 1)
    assert (s["rowsize"], s["table_size"], s["measurement_size_F"], s["energies_size"]) == (704, 8448, 64, 64)
    assert qcg.sizes(2, (6, 6, 6, 3), 65)["table_size"] == 2816
    assert qcg.sizes(2, (6, 6, 6, 3), 65)["energies_size"] == 128


@pytest.mark.parametrize("N", [2, 3])
def test_planes_round_trip(N):
    rng = np.random.default_rng(5)
    V = 48
    links = haar(rng, V, N)
    if N == 2:   # SU(2): [[a, b], [-conj b, conj a]]
        a = links[:, :, 0, 0, 0] + 1j * links[:, :, 0, 0, 1]
        b = links[:, :, 0, 1, 0] + 1j * links[:, :, 0, 1, 1]
        n = np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2)
        a, b = a / n, b / n
        m = np.stack([np.stack([a, b], -1), np.stack([-np.conj(b), np.conj(a)], -1)], -2)
        links = np.stack([m.real, m.imag], axis=-1)
    s = qcg.sizes(N, (4, 4, 3, 1), 1)
    t = qcg.planes_from_links(links, N, s["rowsize"], s["table_size"])
    assert t.shape == (s["table_size"], 4)
    back = qcg.links_from_planes(t, N, V, s["rowsize"])
    # stored rows exact; the rebuilt row is the unitary completion
    assert np.array_equal(back[:, :, : (2 if N == 3 else 1)], links[:, :, : (2 if N == 3 else 1)])
    assert np.max(np.abs(back - links)) < 1e-14
    # the padding between blocks stays zero
    assert not np.any(t[V: s["rowsize"]])


def _write(tmp_path, name, header, links, **kw):
    p = tmp_path / name
    p.write_bytes(qcg.encode(header, links, **kw))
    return p


def test_library_reads_reference_header(tmp_path):
    rng = np.random.default_rng(1)
    dims = (4, 4, 4, 2)
    links = haar(rng, 128, 3)
    h = qcg.make_header(3, dims, 5.9, "single", "XOR128", start=0, randseries=77, prng_counter=45, nav=3,
                        nav_counter=3, niter=2, iter=10, iter_counter=6, nhit=7, wilson_r=3, wilson_t=1,
                        pl_level=2, get_wilson=1, init=1)
    path = _write(tmp_path, "a.qcg", h, links, plq=rng.normal(size=(6, 2)), poly=rng.normal(size=(6, 2)),
                  wilson=rng.normal(size=6))
    p, info, exact = read_qcg_header(path)
    assert not exact
    assert (p.sun, tuple(p.n), p.precision, p.prng, p.randseries, p.nhit) == (3, dims, _abi.FP32,
                                                                             _abi.PRNG_XOR128, 77, 7)
    assert p.beta == 5.9 and p.storage_reals == 18 and p.stream_mode == _abi.STREAM_REFERENCE
    assert p.max_measurements == 10
    assert info == dict(init=1, start=0, prng_counter=45, nav=3, nav_counter=3, niter=2, iter=10, iter_counter=6,
                        wilson_r=3, wilson_t=1, pl_level=2, get_actions=1, get_plaquettes=1, get_wilson=1,
                        fmunu_index1=3, fmunu_index2=8, get_fmunu=0, get_f0mu=0, phi=0.0, omega=0.0)
    # RANLUX luxury and PM codes, SU(2), double precision
    h2 = qcg.make_header(2, (2, 2, 2, 2), 2.3, "double", "RANLUX2", iter=3)
    p2, _, _ = read_qcg_header(_write(tmp_path, "b.qcg", h2, haar(rng, 16, 2)))
    assert (p2.sun, p2.precision, p2.prng, p2.ranlux_luxury, p2.storage_reals) == (2, _abi.FP64, _abi.PRNG_RANLUX,
                                                                                  2, 4)
    h3 = qcg.make_header(3, (2, 2, 2, 2), 6.0, "single", "PM", iter=3)
    assert read_qcg_header(_write(tmp_path, "c.qcg", h3, haar(rng, 16, 3)))[0].prng == _abi.PRNG_PM


@pytest.mark.parametrize("field", ["H", "E"])
def test_fmunu_sections(tmp_path, field):
    """With F_munu the partials and plaquette sections are 7 x longer
    (suncl.cpp:2266-2298); the plaquette section holds the plaquette sums,
    then F for the three planes of the first generator and of the second
    (reduce_measurement_plq_double2, sun_measurements_cl.cl:1312-1354).  An
    E-field run writes get_Fmunu 0 with the longer sections."""
    rng = np.random.default_rng(7)
    dims = (4, 4, 2, 2)
    links = haar(rng, 64, 3)
    h = qcg.make_header(3, dims, 5.9, "double", "XOR128", f_sections=True, iter=5, iter_counter=5,
                        fmunu_index1=2, fmunu_index2=6, get_fmunu=1 if field == "H" else 0)
    plq = rng.normal(size=(5, 2))
    F = rng.normal(size=(5, 2, 3)) + 1j * rng.normal(size=(5, 2, 3))
    data = qcg.encode(h, links, plq=plq, poly=rng.normal(size=(5, 2)), fmunu=F)
    s = qcg.sizes(3, dims, 5)
    assert len(data) == 4 * 64 + 16 * 7 * s["measurement_size_F"] + 16 * s["energies_size"] * (1 + 7) \
        + 16 * s["polyakov_loop_size"] + 32 * s["table_size"]
    d = qcg.decode(data)
    assert np.array_equal(d["plaquettes"][:5], plq) and np.array_equal(d["fmunu"][:5], F)
    # block 1 + 3 g + plane starts at (1 + 3 g + plane) * energies_size double2
    raw = np.frombuffer(data, dtype=np.float64, offset=4 * 64 + 16 * 7 * s["measurement_size_F"]
                        + 16 * s["energies_size"]).reshape(-1, 2)
    en = s["energies_size"]
    assert raw[(1 + 3 * 1 + 2) * en + 4, 0] == F[4, 1, 2].real
    path = tmp_path / "f.qcg"
    path.write_bytes(data)
    _, info, _ = read_qcg_header(path)
    assert (info["fmunu_index1"], info["fmunu_index2"]) == (2, 6)
    assert (info["get_fmunu"], info["get_f0mu"]) == ((1, 0) if field == "H" else (0, 1))


@pytest.mark.parametrize(
    "change, message",
    [
        (lambda w: w.__setitem__(0, 0), "prefix"),
        (lambda w: w.__setitem__(2, qcg.string_word("1.4", 0)), "version"),
        (lambda w: w.__setitem__(qcg.WORD["get_fmunu"], 1), "get_Fmunu but no F_munu sections"),
        (lambda w: w.__setitem__(qcg.WORD["energies_size_F"], 3 * w[qcg.WORD["energies_size"]]), "energies_size_F"),
        (lambda w: w.__setitem__(qcg.WORD["phi_hi"], 0x7FF80000), "PHI / OMEGA not finite"),
        (lambda w: w.__setitem__(42, 6), "multi-device"),
        (lambda w: w.__setitem__(qcg.WORD["generator"], 8), "generator"),
        (lambda w: w.__setitem__(qcg.WORD["nd"], 3), "ND"),
        (lambda w: w.__setitem__(qcg.WORD["precision"], 3), "precision"),
    ],
)
def test_library_rejects(tmp_path, change, message):
    rng = np.random.default_rng(2)
    h = qcg.make_header(3, (4, 4, 4, 2), 5.9, "single", "XOR128", iter=4)
    change(h)
    path = _write(tmp_path, "x.qcg", h, haar(rng, 128, 3))
    with pytest.raises(QCDError, match=message):
        read_qcg_header(path)


def test_library_rejects_truncated(tmp_path):
    rng = np.random.default_rng(3)
    h = qcg.make_header(3, (4, 4, 4, 2), 5.9, "single", "XOR128", iter=4)
    data = qcg.encode(h, haar(rng, 128, 3))
    path = tmp_path / "t.qcg"
    path.write_bytes(data[:-16])
    with pytest.raises(QCDError, match="truncated"):
        read_qcg_header(path)
    with pytest.raises(QCDError, match="cannot open"):
        read_qcg_header(tmp_path / "missing.qcg")


@pytest.mark.parametrize("n", [2**64 - 1, 2**63, 1 << 40, 17])
def test_library_ignores_bad_extension_length(tmp_path, n):
    """An extension block whose byte count does not fit in what follows it
    (a huge count would wrap an unsigned bound test) is treated as absent:
    the file reads as a reference file, no state image."""
    import struct

    rng = np.random.default_rng(4)
    h = qcg.make_header(3, (4, 4, 4, 2), 5.9, "single", "XOR128", iter=4)
    data = qcg.encode(h, haar(rng, 128, 3)) + b"QCDGPUX1" + struct.pack("<Q", n) + b"\0" * 8
    path = tmp_path / "e.qcg"
    path.write_bytes(data)
    p, _, exact = read_qcg_header(path)
    assert not exact and p.sun == 3
