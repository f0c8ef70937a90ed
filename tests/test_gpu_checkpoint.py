"""Checkpoint / resume on the device (model::lattice_save_state /
lattice_load_state, suncl.cpp:1857-2079; generator replay suncl.cpp:5349-5351).

Bars: a run continued from a state image or a .qcg file is bit-identical to
the uninterrupted run (links, records, generator states); replayed streams
are bit-identical to the streams of the run they replay; the .qcg sections
decode with the oracle's restatement of the reference layout (oracle/qcg.py)
to exactly the lattice's own links (FP32 tables as float32) and records."""
import os
import subprocess

import numpy as np
import pytest

import qcg
import qcdgpu_amd as Q

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "qcdgpu_amd", "qcdgpu_hip")
STATE_HEAD = 128   # bytes before the table in a state image


def _case(msg):
    print("[case]", msg, flush=True)


def _lat(sun=3, dims=(8, 8, 8, 4), prec="fp32", storage=None, prng="XOR128", stream="reference", beta=5.9,
         seed=54, max_measurements=16):
    lux = 3
    if prng.startswith("RANLUX"):
        lux = int(prng[6:] or 3)
        prng = "RANLUX"
    return Q.Lattice(sun=sun, dims=dims, beta=beta, precision=prec, storage_reals=storage, nhit=10, prng=prng,
                     ranlux_luxury=lux, stream=stream, randseries=seed, max_measurements=max_measurements)


def _start(lat, start):
    if start == "hot":
        lat.prng_init()
        lat.init("hot")
    else:
        lat.init(start)
        if start == "cold":
            lat.prng_init()
    lat.reset_measurements()


def _cycles(lat, n, n_or=1, wilson=True):
    for _ in range(n):
        lat.sweep("hb")
        for _ in range(n_or):
            lat.sweep("or")
        lat.measure(wilson=wilson)


def _same_run(a, b):
    assert np.array_equal(a.export_links(), b.export_links())
    pa, qa = a.raw_measurements()
    pb, qb = b.raw_measurements()
    assert np.array_equal(pa, pb) and np.array_equal(qa, qb)
    assert np.array_equal(a.wilson_measurements(raw=True), b.wilson_measurements(raw=True))
    assert a.prng_counter() == b.prng_counter()
    assert np.array_equal(a.save_state(), b.save_state())


def _rng_slice(lat, img):
    m = lat.params.max_measurements
    slices = 2 * 8 * (lat.dims[0] + lat.dims[1] + lat.dims[2]) * m   # Polyakov slice records
    fmunu = 12 * 8 * m                                                # F_munu records
    actions = slices                                                  # action slice records
    n = img.size - STATE_HEAD - lat.table_bytes() - 7 * 8 * m - slices - fmunu - actions
    return slice(STATE_HEAD + lat.table_bytes(), STATE_HEAD + lat.table_bytes() + n)


CONFIGS = [
    dict(sun=3, prec="fp32", storage=18, prng="XOR128"),
    dict(sun=3, prec="fp64", storage=12, prng="RANLUX3"),
    dict(sun=2, prec="fp64", storage=4, prng="PM"),
    dict(sun=3, prec="fp32", storage=12, prng="XOR128", stream="fast"),
]
IDS = ["su3-fp32-18-xor", "su3-fp64-12-ranlux3", "su2-fp64-pm", "su3-fp32-12-xor-fast"]


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_state_image_continues_bit_identically(cfg):
    _case("state image %s" % cfg)
    with _lat(**cfg) as a, _lat(**cfg) as b:
        a.set_wilson(3, 2)
        _start(a, "hot")
        a.measure(wilson=True)
        _cycles(a, 2)
        b.load_state(a.save_state())
        _same_run(a, b)
        _cycles(a, 3)
        _cycles(b, 3)
        _same_run(a, b)


def test_state_image_refuses_another_lattice():
    _case("state image mismatch")
    with _lat() as a, _lat(beta=6.0) as b, _lat(dims=(8, 8, 8, 6)) as c:
        _start(a, "cold")
        img = a.save_state()
        with pytest.raises(Q.QCDError, match="beta"):
            b.load_state(img)
        with pytest.raises(Q.QCDError, match="bytes"):
            c.load_state(img)


@pytest.mark.parametrize("prng", ["XOR128", "PM", "RANLUX1"])
@pytest.mark.parametrize("start", ["cold", "hot", "random"])
def test_prng_replay_matches_run(prng, start):
    """Streams re-seeded and advanced to a run's production counter equal
    that run's streams; 6^3 x 4 has 432 links per half-step, so the last
    64-link unit carries idle streams (the hot start leaves them alone, the
    half-steps advance them)."""
    _case("replay %s %s" % (prng, start))
    dims = (6, 6, 6, 4)
    with _lat(dims=dims, prng=prng) as a, _lat(dims=dims, prng=prng) as b:
        _start(a, start)
        for _ in range(3):
            a.sweep("hb")
            a.sweep("or")
        a.halfstep(1, 0, "hb")
        c = a.prng_counter()
        assert c == 1 + (0 if start == "cold" else 4) + 8 * 3 + 1
        b.prng_replay(c, start)
        assert b.prng_counter() == c
        ia, ib = a.save_state(), b.save_state()
        s = _rng_slice(a, ia)
        assert s.stop > s.start
        assert np.array_equal(ia[s], ib[s])


def test_prng_counter_accounting():
    _case("counter accounting")
    with _lat() as a:
        _start(a, "cold")
        assert a.prng_counter() == 1
        a.sweep("hb")
        a.sweep("or")
        assert a.prng_counter() == 9
        a.capture_cycle(n_hb=2, n_or=1)
        assert a.prng_counter() == 9   # capture runs nothing
        a.launch_cycle(3)
        a.sync()
        assert a.prng_counter() == 9 + 3 * 16
        a.time_cycle(2)
        assert a.prng_counter() == 9 + 5 * 16
    with _lat() as g:
        _start(g, "gid")
        g.sweep("hb")
        assert g.prng_counter() == 0


QCG_CASES = [
    dict(sun=3, prec="fp32", storage=18, prng="XOR128"),
    dict(sun=3, prec="fp64", storage=18, prng="RANLUX2"),
    dict(sun=2, prec="fp32", storage=4, prng="PM"),
]


@pytest.mark.parametrize("cfg", QCG_CASES, ids=["su3-fp32", "su3-fp64", "su2-fp32"])
def test_qcg_sections_decode_with_reference_layout(tmp_path, cfg):
    _case("qcg layout %s" % cfg)
    with _lat(**cfg) as a:
        a.set_wilson(2, 2)
        _start(a, "hot")
        a.measure(wilson=True)
        _cycles(a, 2)
        path = tmp_path / "s.qcg"
        a.save_qcg(path, start="hot", iter=10, nav=0, niter=1, iter_counter=3, get_wilson=1, wilson_r=2,
                   wilson_t=2, pl_level=3)
        assert not os.path.exists(str(path) + ".tmp")
        d = qcg.decode(path.read_bytes())
        h, s = d["header"], d["sizes"]
        assert h["prefix"] == (qcg.string_word("QCDGPU", 0), qcg.string_word("QCDGPU", 4))
        assert h["group"] == cfg["sun"] and h["full"] == (8, 8, 8, 4) == h["domain"]
        assert h["beta"] == 5.9 and h["prng_counter"] == a.prng_counter() == 5 + 16
        assert h["precision"] == (1 if cfg["prec"] == "fp32" else 2) and h["iter_counter"] == 3
        assert h["measurement_size_F"] == s["measurement_size_F"] and h["energies_size"] == 64
        links = a.export_links()
        want = qcg.planes_from_links(links, cfg["sun"], s["rowsize"], s["table_size"])
        if cfg["prec"] == "fp32":
            want = want.astype(np.float32)
        assert np.array_equal(d["table"], want)
        plq, poly = a.raw_measurements()
        assert np.array_equal(d["plaquettes"][:3], plq) and not np.any(d["plaquettes"][3:])
        assert np.array_equal(d["polyakov"][:3], poly[:, :2])
        assert np.array_equal(d["wilson"][:3], a.wilson_measurements(raw=True))
        N, V = cfg["sun"], 8 * 8 * 8 * 4
        assert np.allclose(d["actions"][:3], 5.9 / N * (3 * N * V - plq), rtol=1e-15, atol=1e-9)
        tail = d["tail"]
        assert tail[:8] == b"QCDGPUX1"
        assert int.from_bytes(tail[8:16], "little") == len(tail) - 16 == a.save_state().size


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_qcg_resume_exact(tmp_path, cfg):
    _case("qcg resume exact %s" % cfg)
    path = tmp_path / "r.qcg"
    with _lat(**cfg) as a:
        _start(a, "cold")
        a.measure(wilson=True)
        _cycles(a, 2)
        a.save_qcg(path, iter=16, iter_counter=3, get_wilson=1, wilson_r=7, wilson_t=2)
        b, info = Q.Lattice.from_qcg(path)
        with b:
            assert info["iter_counter"] == 3 and info["iter"] == 16
            _same_run(a, b)
            _cycles(a, 3)
            _cycles(b, 3)
            _same_run(a, b)


FALLBACK = [
    (dict(sun=3, prec="fp32", storage=12, prng="XOR128"), "random"),
    (dict(sun=2, prec="fp64", storage=4, prng="PM"), "cold"),
    (dict(sun=3, prec="fp64", storage=12, prng="RANLUX3"), "random"),
]


@pytest.mark.parametrize("cfg,start", FALLBACK, ids=["su3-fp32-12-xor-hot", "su2-fp64-pm", "su3-fp64-ranlux3"])
def test_qcg_resume_from_reference_sections(tmp_path, cfg, start):
    """A file with only the reference's sections (the extension cut off):
    table, plaquette / Polyakov (Re, Im) / Wilson sums and the production
    counter, with the streams replayed (a hot, ints = 0, file is the
    reference's random start, replayed with the hot runs' produce() size).
    Two-row storage holds exactly what the file holds, so the continued run
    is bit-identical."""
    _case("qcg resume sections %s %s" % (cfg, start))
    full, cut = tmp_path / "f.qcg", tmp_path / "c.qcg"
    with _lat(**cfg) as a, _lat(**cfg) as b:
        a.set_wilson(2, 2)
        _start(a, start)
        a.measure(wilson=True)
        _cycles(a, 2)
        a.save_qcg(full, start=start, iter=16, iter_counter=3, get_wilson=1, wilson_r=2, wilson_t=2)
        d = qcg.decode(full.read_bytes())
        cut.write_bytes(full.read_bytes()[: d["table_end"]])
        assert not Q.read_qcg_header(cut)[2]
        info = b.load_qcg(cut)
        assert info["iter_counter"] == 3 and b.prng_counter() == a.prng_counter()
        assert np.array_equal(a.export_links(), b.export_links())
        _cycles(a, 3)
        _cycles(b, 3)
        assert np.array_equal(a.export_links(), b.export_links())
        pa, qa = a.raw_measurements()
        pb, qb = b.raw_measurements()
        assert np.array_equal(pa, pb)
        assert np.array_equal(qa[:, :2], qb[:, :2])
        assert np.array_equal(qa[3:], qb[3:]) and not np.any(qb[:3, 2:])   # P2, P4 are not in the file
        assert np.array_equal(a.wilson_measurements(raw=True), b.wilson_measurements(raw=True))


@pytest.mark.parametrize("start", ["hot", "random"])
def test_qcg_sections_replay_follows_the_files_start(tmp_path, start):
    """A file whose image does not fit the lattice (another storage) resumes
    from its sections; the image still tells how a hot (ints = 0) run began:
    the Haar start (cold-size productions) or the random start (hot-size
    productions), and the streams are replayed accordingly."""
    _case("qcg sections replay %s" % start)
    path = tmp_path / "s.qcg"
    with _lat(prng="PM", prec="fp64", storage=18) as a, _lat(prng="PM", prec="fp64", storage=12) as b:
        _start(a, start)
        _cycles(a, 2, wilson=False)
        a.save_qcg(path, start=start, iter=16, iter_counter=3)
        b.load_qcg(path)
        assert b.prng_counter() == a.prng_counter()
        ia, ib = a.save_state(), b.save_state()
        assert np.array_equal(ia[_rng_slice(a, ia)], ib[_rng_slice(b, ib)])


def _rows(path):
    rows = []
    for line in open(path):
        parts = line.split()
        if len(parts) == 28 and parts[0].isdigit():   # 11 columns and 16 of F_munu (on by default)
            rows.append(parts[1:])
    return rows


def _run(dirpath, text):
    f = dirpath / "init.dat"
    f.write_text(text + "OUTPUTPATH = %s/\n" % dirpath)
    return subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=180)


def test_host_stop_and_resume(tmp_path):
    """The host program stopped inside an iteration (work after the last
    checkpoint lost), then continued from its newest .qcg with INIT = 0,
    writes the results file of the uninterrupted run, digit for digit."""
    _case("host stop and resume")
    base = ("GROUP = 3\nLS = 8\nLT = 4\nITER = 12\nNITER = 2\nNAV = 3\nBETA = 5.9\nINTS = 0\nPRNG = XOR128\n"
            "RANDSERIES = 5\nPRECISION = SINGLE\nNOR = 1\nGETWILSON = 1\nWILSONR = 2\nWILSONT = 2\n")
    da, db, dc = (tmp_path / k for k in "abc")
    for d in (da, db, dc):
        d.mkdir()
    ra = _run(da, base)
    assert ra.returncode == 0, ra.stderr
    rb = _run(db, base + "SAVESTATE = 1\nSAVEEVERY = 0\nSTOPAFTER = 7\n")
    assert rb.returncode == 3, rb.stdout + rb.stderr
    states = sorted((p for p in os.listdir(db) if p.endswith(".qcg")), key=lambda p: os.path.getmtime(db / p))
    assert states and not [p for p in os.listdir(db) if p.endswith(".tmp")]
    newest = db / states[-1]
    p, info, exact = Q.read_qcg_header(newest)
    assert exact and info["iter_counter"] == 8 and info["nav_counter"] == 3 and info["start"] == 0
    # keys the file does not hold (NOR) still come from init.dat
    rc = _run(dc, base + "INIT = 0\nSTATEFILE = %s\n" % newest)
    assert rc.returncode == 0, rc.stderr
    out_a = [x for x in os.listdir(da) if x.endswith(".txt")]
    out_c = [x for x in os.listdir(dc) if x.endswith(".txt")]
    assert len(out_a) == 1 and len(out_c) == 1
    rows_a, rows_c = _rows(da / out_a[0]), _rows(dc / out_c[0])
    assert len(rows_a) == 12 and rows_a == rows_c


@pytest.mark.parametrize("field", ["H", "E"])
def test_qcg_fmunu_sections(tmp_path, field):
    """F_munu records travel in the reference's longer plaquette section
    (7 x energies_size: plaquettes, then the planes of each generator) and
    in the exact image; a load restores them and the generators."""
    _case("qcg fmunu %s" % field)
    cfg = dict(sun=3, prec="fp32", storage=12, prng="XOR128")
    full, cut = tmp_path / "f.qcg", tmp_path / "c.qcg"
    with _lat(**cfg) as a, _lat(**cfg) as b:
        _start(a, "hot")
        a.set_fmunu(2, 6)
        a.measure(fmunu=field)
        _cycles(a, 1)
        a.measure(fmunu=field)
        F = a.fmunu(raw=True)
        assert F.shape[0] >= 2 and np.any(F[-1])
        a.save_qcg(full, start="hot", iter=16, fmunu_index1=2, fmunu_index2=6,
                   get_fmunu=int(field == "H"), get_f0mu=int(field == "E"))
        d = qcg.decode(full.read_bytes())
        h = d["header"]
        assert h["energies_size_F"] == 7 * h["energies_size"] and h["get_fmunu"] == int(field == "H")
        assert (h["fmunu_index1"], h["fmunu_index2"]) == (2, 6)
        n = F.shape[0]
        plq, _ = a.raw_measurements()
        assert np.array_equal(d["plaquettes"][:n], plq[:n])
        assert np.array_equal(d["fmunu"][:n], F) and not np.any(d["fmunu"][n:])
        cut.write_bytes(full.read_bytes()[: d["table_end"]])
        info = b.load_qcg(cut)
        assert (info["get_fmunu"], info["get_f0mu"]) == (int(field == "H"), int(field == "E"))
        assert np.array_equal(b.fmunu(raw=True), F)
        c, _ = Q.Lattice.from_qcg(full)        # exact image
        with c:
            assert np.array_equal(c.fmunu(raw=True), F)
            for lat in (a, b, c):               # generators restored from the header
                lat.measure(fmunu=field)
            assert np.array_equal(b.fmunu(raw=True)[-1], a.fmunu(raw=True)[-1])
            assert np.array_equal(c.fmunu(raw=True)[-1], a.fmunu(raw=True)[-1])


def test_qcg_twist(tmp_path):
    """PHI / OMEGA travel in header words 21-24 (lattice_make_bin_header,
    suncl.cpp:1967-1970) and are applied on load, by the exact image and by
    the reference sections; the continued runs stay identical."""
    _case("qcg twist")
    cfg = dict(sun=3, prec="fp32", storage=12, prng="XOR128")
    full, cut = tmp_path / "t.qcg", tmp_path / "c.qcg"
    with _lat(**cfg) as a, _lat(**cfg) as b:
        a.set_twist(3.0, 0.5)
        _start(a, "cold")
        _cycles(a, 2)
        a.save_qcg(full, iter=16)
        w = np.frombuffer(full.read_bytes()[:256], dtype=np.uint32)
        assert w[21:25].view(np.float64).tolist() == [3.0, 0.5]
        assert Q.read_qcg_header(full)[1]["phi"] == 3.0
        d = qcg.decode(full.read_bytes())
        cut.write_bytes(full.read_bytes()[: d["table_end"]])
        b.load_qcg(cut)
        c, _ = Q.Lattice.from_qcg(full)
        with c:
            assert b.twist() == (3.0, 0.5) == c.twist()
            for lat in (a, b, c):
                _cycles(lat, 2)
            assert np.array_equal(a.export_links(), b.export_links())
            assert np.array_equal(a.export_links(), c.export_links())
            pa, _ = a.raw_measurements()
            assert np.array_equal(pa, c.raw_measurements()[0])
            assert np.array_equal(pa[-2:], b.raw_measurements()[0][-2:])
