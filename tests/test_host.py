"""The command-line host program (qcdgpu_amd/qcdgpu_hip): init.dat parsing
and defaults on the CPU; a run through the C ABI on the GPU, checked against
the Python driver on the same parameters."""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "qcdgpu_amd", "qcdgpu_hip")


def _params(text, tmp_path):
    f = tmp_path / "init.dat"
    f.write_text(text)
    r = subprocess.run([BIN, "--print-params", str(f)], capture_output=True, text=True)
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    return r.returncode, kv


def test_host_built():
    assert os.access(BIN, os.X_OK), "build with __graft_entry__.build()"


def test_defaults_are_the_reference_defaults():
    """SU(N)/QCDGPU.cpp:98-198: SU(3) 32^4, beta 6.5, cold start, RANLUX3,
    random series 54, NHIT 10, ITER 2, NITER 1, NAV 0, double precision."""
    r = subprocess.run([BIN, "--print-params"], capture_output=True, text=True)
    assert r.returncode == 0
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines())
    assert kv["GROUP"] == "3" and [kv["L%d" % k] for k in range(1, 5)] == ["32"] * 4
    assert float(kv["BETA"]) == 6.5 and kv["INTS"] == "1" and kv["PRNG"] == "RANLUX3"
    assert kv["RANDSERIES"] == "54" and kv["NHIT"] == "10" and kv["ITER"] == "2"
    assert kv["NITER"] == "1" and kv["NAV"] == "0" and kv["PRECISION"] == "DOUBLE"


def test_reference_init_syntax(tmp_path):
    """Tab-indented KEY = VALUE lines, exponent floats, '#' comments, LS/LT
    then L1 (a later key overrides), and the INTS start codes (0 hot)."""
    rc, kv = _params(
        "\tGROUP = 2\n\tND = 4\n\tLT = 8\n\tLS = 12\n\tL1 = 16   # x\n\tITER = 50\n\tNITER = 10\n"
        "\tBETA = 2.3000000000000e+00\n\tNAV = 5\n\tINTS = 0\n\tPRNG = RANLUX2\n\tRANDSERIES = 7\n"
        "\tWILSONR = 10\n\tPRECISION = SINGLE\n\tNOR = 3\n",
        tmp_path,
    )
    assert rc == 0
    assert (kv["L1"], kv["L2"], kv["L3"], kv["L4"]) == ("16", "12", "12", "8")
    assert kv["GROUP"] == "2" and float(kv["BETA"]) == 2.3 and kv["INTS"] == "0" and kv["HAAR"] == "0"
    assert kv["PRNG"] == "RANLUX2" and kv["RANDSERIES"] == "7" and kv["ITER"] == "50"
    assert kv["PRECISION"] == "SINGLE" and kv["NOR"] == "3" and kv["STORAGE"] == "4"


def test_haar_key(tmp_path):
    """INTS = 0 is the reference's hot start; HAAR = 1 (a key of this build)
    asks for the Haar start instead."""
    rc, kv = _params("INTS = 0\nHAAR = 1\n", tmp_path)
    assert rc == 0 and kv["INTS"] == "0" and kv["HAAR"] == "1"


@pytest.mark.parametrize(
    "line", ["PRNG = RANMAR", "ND = 3", "GROUP = 4", "PHI = nan",
             "GETWILSON = 1\nWILSONR = 0"]
)
def test_unsupported_features_rejected(tmp_path, line):
    rc, kv = _params(line + "\n", tmp_path)
    assert rc == 2 and "UNSUPPORTED" in kv


@pytest.mark.parametrize(
    "text, want",
    [
        ("", ("H", "3", "8")),                              # get_Fmunu is on by default (suncl.cpp:761)
        ("F0MU = 1\n", ("E", "3", "8")),
        ("FMUNU = 0\nTURNOFFFMUNU = 1\n", ("OFF", "3", "8")),
        ("FMUNU1 = 1\nFMUNU6 = 1\n", ("H", "1", "6")),
        ("FMUNU1 = 1\nFMUNU4 = 1\nFMUNU7 = 1\nFMUNU5 = 1\n", ("H", "4", "5")),   # a later key replaces
        ("GROUP = 2\nFMUNU1 = 1\nFMUNU6 = 1\n", ("H", "1", "2")),   # SU(2): sigma_1 or sigma_3, sigma_2
        ("GROUP = 2\nFMUNU2 = 1\n", ("H", "3", "2")),
    ],
)
def test_fmunu_keys(tmp_path, text, want):
    """model::lattice_get_init_file F_munu keys (suncl.cpp:412-453) act by
    presence; Fmunu_defaults picks the generators (suncl.cpp:2093-2113)."""
    rc, kv = _params(text, tmp_path)
    assert rc == 0
    assert (kv["FMUNU"], kv["FMUNU_INDEX1"], kv["FMUNU_INDEX2"]) == want


def test_twist_keys(tmp_path):
    """PHI / OMEGA (suncl.cpp:372-373) are read as numbers; the reference's
    own init.dat sets PHI = 3."""
    rc, kv = _params("PHI = 3.0000000000000e+00\nOMEGA = -0.25\n", tmp_path)
    assert rc == 0 and float(kv["PHI"]) == 3.0 and float(kv["OMEGA"]) == -0.25


def test_bad_precision_value(tmp_path):
    rc, _ = _params("PRECISION = HALF\n", tmp_path)
    assert rc == 1


@pytest.mark.gpu
@pytest.mark.parametrize("prng,ints,fkeys", [("XOR128", 1, ""), ("RANLUX3", 0, "F0MU = 1\nFMUNU1 = 1\nFMUNU7 = 1\n"),
                                             ("PM", 0, "PHI = 3.0\nOMEGA = 0.5\nACTIONSDIFF = 1\n"),
                                             ("XOR128", 0, "HAAR = 1\n")])
def test_host_run_matches_python_driver(tmp_path, prng, ints, fkeys):
    """qcdgpu_hip on a small lattice writes the same plaquette, Wilson loop,
    Polyakov and F_munu series as qcdgpu_amd.Lattice.simulate with the same
    parameters (F_munu is on by default, as in the reference)."""
    import qcdgpu_amd as Q

    print("[case] host run %s ints=%d %s" % (prng, ints, fkeys.split("\n")[0]), flush=True)
    f = tmp_path / "init.dat"
    f.write_text(
        "GROUP = 3\nLS = 8\nLT = 4\nITER = 4\nNITER = 2\nNAV = 1\nBETA = 5.9\nINTS = %d\nPRNG = %s\n"
        "RANDSERIES = 11\nPRECISION = DOUBLE\nNOR = 2\nGETWILSON = 1\nWILSONR = 2\nWILSONT = 2\n"
        "OUTPUTPATH = %s/\n%s" % (ints, prng, tmp_path, fkeys)
    )
    efield = "F0MU" in fkeys
    g1, g2 = (1, 7) if "FMUNU1" in fkeys else (3, 8)
    twist = (3.0, 0.5) if "PHI" in fkeys else (0.0, 0.0)
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the end-of-run check of the final configuration on the host
    assert "Verification successfully passed!" in r.stdout and "test failed" not in r.stdout, r.stdout
    out = [p for p in os.listdir(tmp_path) if p.startswith("su3-") and p.endswith(".txt")]
    assert len(out) == 1
    rows, pl_rows, s_rows, section, fields = [], {}, {}, None, None
    lines = open(tmp_path / out[0]).read().splitlines()
    assert " *** Verification successfully passed! *************" in lines
    assert any(ln.startswith(" CPU last Wilson_loop") for ln in lines)
    for k, line in enumerate(lines):
        parts = line.split()
        if line.strip() == "Data fields:":
            fields = lines[k + 1].split()[1:]
        if len(parts) == 28 and parts[0].isdigit():
            rows.append([float(x) for x in parts[1:]])
        if parts and parts[0] in ("X", "Y", "Z") and len(parts) == 9:
            section = parts[0]
        elif parts and parts[0] in ("PL", "PL_im") and section:
            pl_rows[section, parts[0]] = np.array([float(x) for x in parts[1:]])
        elif parts and parts[0] in ("S", "S_variance") and section:
            s_rows[section, parts[0]] = np.array([float(x) for x in parts[1:]])
    rows = np.array(rows)
    assert rows.shape == (4, 27)   # Plq x3, Wilson_loop, S x3, Polyakov x4, F_munu x16
    planes = ("xt", "yt", "zt") if efield else ("xy", "xz", "yz")
    assert any(ln.startswith(" PHI") and float(ln.split(":")[1]) == twist[0] for ln in lines)
    assert fields[11] == "Fmunu_%s_%d_re" % (planes[0], g1) and fields[17] == "Fmunu_%s_%d_re" % (planes[0], g2)
    assert fields[26] == "Fmunu_abs_%d_im" % g2
    assert len(pl_rows) == 6        # PL_LEVEL 3 (the default): the differentiated Polyakov loop block
    lux = int(prng[-1]) if prng.startswith("RANLUX") else 3
    name = "RANLUX" if prng.startswith("RANLUX") else prng
    with Q.Lattice(sun=3, dims=(8, 8, 8, 4), beta=5.9, precision="fp64", prng=name, ranlux_luxury=lux,
                   randseries=11, max_measurements=8) as lat:
        lat.set_wilson(2, 2)
        lat.set_fmunu(g1, g2)
        lat.set_twist(*twist)
        start = "hot" if "HAAR" in fkeys else {0: "random", 1: "cold"}[ints]
        m = lat.simulate(nav=1, iters=4, niter=2, n_or=2, start=start, wilson=True,
                         slices=True, fmunu="E" if efield else "H", action_slices="ACTIONSDIFF" in fkeys)
        asl = lat.action_slices()
        wl = lat.wilson_measurements()
        sl = lat.polyakov_slices()
        F = lat.fmunu()
    assert np.allclose(rows[:, 0], m["plq_spat"], rtol=1e-13, atol=0)
    assert np.allclose(rows[:, 1], m["plq_temp"], rtol=1e-13, atol=0)
    assert np.allclose(rows[:, 3], wl, rtol=1e-12, atol=1e-15)
    assert np.allclose(rows[:, 7], m["polyakov_re"], rtol=1e-12, atol=1e-15)
    for d in "xyz":   # slice means without measurement 0, as lattice_data_analysis
        assert np.allclose(pl_rows[d.upper(), "PL"], sl[d][1:].real.mean(axis=0), rtol=1e-11, atol=1e-14)
        assert np.allclose(pl_rows[d.upper(), "PL_im"], sl[d][1:].imag.mean(axis=0), rtol=1e-11, atol=1e-14)
    for gi in range(2):
        for pl in range(3):
            assert np.allclose(rows[:, 11 + 6 * gi + 2 * pl], F[:, gi, pl].real, rtol=1e-12, atol=1e-15)
            assert np.allclose(rows[:, 12 + 6 * gi + 2 * pl], F[:, gi, pl].imag, rtol=1e-12, atol=1e-15)
        # |F| = sqrt of the sum over planes of squares (lattice_data_analysis_joint3)
        assert np.allclose(rows[:, 23 + 2 * gi], np.sqrt((F[:, gi].real ** 2).sum(axis=1)), rtol=1e-12, atol=1e-15)
        assert np.allclose(rows[:, 24 + 2 * gi], np.sqrt((F[:, gi].imag ** 2).sum(axis=1)), rtol=1e-12, atol=1e-15)
    if "ACTIONSDIFF" in fkeys:   # S_total = (S_s + S_t) / 2 per slice, mean and variance without measurement 0
        assert len(s_rows) == 6
        for d in "xyz":
            tot = 0.5 * (asl[d][1:, :, 0] + asl[d][1:, :, 1])
            assert np.allclose(s_rows[d.upper(), "S"], tot.mean(axis=0), rtol=1e-11, atol=1e-14)
            assert np.allclose(s_rows[d.upper(), "S_variance"], tot.var(axis=0), rtol=1e-8, atol=1e-18)
    else:
        assert not s_rows


GOLDEN = os.path.join(REPO, "tests", "golden")


def test_reference_init_dat_accepted(tmp_path):
    """The reference's own SU(N)/init.dat (a copy in tests/golden/): SU(2)
    16^4, beta 3.5, RANLUX3, Wilson loop 10 x 3, PL_LEVEL 3, PHI = 3 (the
    twisted boundary) and FMUNU -- every key it sets is taken."""
    r = subprocess.run([BIN, "--print-params", os.path.join(GOLDEN, "reference_init.dat")], capture_output=True,
                       text=True)
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    assert r.returncode == 0 and "UNSUPPORTED" not in kv, r.stdout
    assert kv["GROUP"] == "2" and [kv["L%d" % k] for k in range(1, 5)] == ["16"] * 4
    assert float(kv["BETA"]) == 3.5 and kv["PRNG"] == "RANLUX3" and kv["ITER"] == "50" and kv["NAV"] == "50"
    assert (kv["WILSONR"], kv["WILSONT"], kv["PL_LEVEL"]) == ("10", "3", "3")
    assert float(kv["PHI"]) == 3.0 and kv["FMUNU"] == "H" and kv["GETWILSON"] == "1"


@pytest.mark.gpu
def test_reference_init_dat_runs(tmp_path):
    """The reference's init.dat, run as it stands (540 sweeps of SU(2) 16^4
    with the twist, Wilson loop, Polyakov slices and F_munu), only
    OUTPUTPATH pointed at a writable folder (a later key replaces an earlier
    one, as in the reference's parser)."""
    print("[case] reference init.dat", flush=True)
    f = tmp_path / "init.dat"
    f.write_text(open(os.path.join(GOLDEN, "reference_init.dat")).read() + "\n\tOUTPUTPATH = %s/\n" % tmp_path)
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = [p for p in os.listdir(tmp_path) if p.startswith("su2-") and p.endswith(".txt")]
    assert len(out) == 1
    text = open(tmp_path / out[0]).read()
    rows = [ln.split() for ln in text.splitlines() if ln.split() and ln.split()[0].isdigit() and len(ln.split()) == 28]
    assert len(rows) == 50
    a = np.array([[float(x) for x in r_[1:]] for r_ in rows])
    assert np.all(np.isfinite(a))
    assert np.all(np.abs(a[:, :3]) <= 2.0)                       # Plq_spat / temp / total: mean ReTr U_p, |.| <= N
    assert 1.0 < a[1:, 2].mean() < 2.0                           # ordered phase at beta = 3.5
    phi = [ln for ln in text.splitlines() if ln.startswith(" PHI")]
    assert len(phi) == 1 and float(phi[0].split(":")[1]) == 3.0
    assert "Differentiated Polyakov loop data" in text and "FMUNU(3, 2)" in text
    assert " *** Verification successfully passed! *************" in text and "Verification successfully passed!" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [18, 12])
def test_host_run_single_precision_check(tmp_path, storage):
    """A single-precision run (FP32 links, FP64 sums): the device's last
    records agree with the host's FP64 measurement of the exported final
    configuration within 5e-5 relative."""
    print("[case] host run single precision storage=%d" % storage, flush=True)
    f = tmp_path / "init.dat"
    f.write_text("GROUP = 3\nLS = 8\nLT = 4\nITER = 3\nNITER = 2\nBETA = 5.9\nINTS = 1\nPRNG = XOR128\n"
                 "PRECISION = SINGLE\nSTORAGE = %d\nNOR = 3\nGETWILSON = 1\nWILSONR = 3\nWILSONT = 2\n"
                 "OUTPUTPATH = %s/\n" % (storage, tmp_path))
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Verification successfully passed!" in r.stdout and "test failed" not in r.stdout, r.stdout


# ---------------------------------------------------------------- end-of-run check
def _device_like_last(O, L, links, sun, dims, beta, wilson=(2, 2), efield=False, gens=(3, 8)):
    """The last values of the results-file columns for `links`, from the
    oracle's sums with the host program's normalisations."""
    V = float(np.prod(dims))
    Vs, N = V / dims[3], float(sun)
    p = O.plaquette(L, links)
    ps, pt = p[0] / (3 * V), p[1] / (3 * V)
    bet = beta / N
    pl = O.polyakov(L, links)
    last = {"Plq_spat": ps, "Plq_temp": pt, "Plq_total": 0.5 * (ps + pt), "S_spat": bet * (N - ps),
            "S_temp": bet * (N - pt), "S_total": 0.5 * (bet * (N - ps) + bet * (N - pt)),
            "Polyakov_loop": pl[0] / (Vs * N), "Polyakov_loop_im": pl[1] / (Vs * N),
            "Polyakov_loop_P2": pl[2] / (Vs * N * N), "Polyakov_loop_P4": pl[3] / (Vs * N ** 4),
            "Wilson_loop": O.wilson(L, links, *wilson) / (3 * V)}
    f = O.fmunu(L, links, efield=efield, gen1=gens[0], gen2=gens[1]) / V
    names = ("xt", "yt", "zt") if efield else ("xy", "xz", "yz")
    for g in range(2):
        for pl_i in range(3):
            last["Fmunu_%s_%d_re" % (names[pl_i], gens[g])] = f[g, pl_i].real
            last["Fmunu_%s_%d_im" % (names[pl_i], gens[g])] = f[g, pl_i].imag
        last["Fmunu_abs_%d_re" % gens[g]] = np.sqrt((f[g].real ** 2).sum())
        last["Fmunu_abs_%d_im" % gens[g]] = np.sqrt((f[g].imag ** 2).sum())
    return last


def _verify(tmp_path, init, links, last):
    (tmp_path / "init.dat").write_text(init)
    np.ascontiguousarray(links, dtype=np.float64).tofile(tmp_path / "links.f64")
    (tmp_path / "last.txt").write_text("".join("%s %.17g\n" % kv for kv in last.items()))
    return subprocess.run([BIN, "--verify", str(tmp_path / "init.dat"), str(tmp_path / "links.f64"),
                           str(tmp_path / "last.txt")], capture_output=True, text=True)


@pytest.mark.parametrize("sun,beta,prec", [(3, 5.9, "DOUBLE"), (3, 5.9, "SINGLE"), (2, 2.3, "DOUBLE")])
def test_end_of_run_check(O, tmp_path, sun, beta, prec):
    """The host's own FP64 measurement of a configuration agrees with the
    oracle's (plaquette, action, Polyakov loop, Wilson loop, F_munu): the
    check passes and prints every CPU / GPU last pair; a record off by 1e-4
    relative fails it at both precisions, and one off by 1e-6 only at double
    precision (tolerances 5e-5 and 1e-11, data_analysis.cpp:40-43)."""
    dims = (4, 4, 2, 4)
    L = O.lattice(sun, dims, beta, 10)
    a = O.cold(L)
    rng = O.SeqRng("RANLUX", 5, 223)
    for _ in range(3):
        O.sweep(L, a, O.HB, rng)
    init = ("GROUP = %d\nL1 = %d\nL2 = %d\nL3 = %d\nL4 = %d\nBETA = %r\nPRECISION = %s\nGETWILSON = 1\n"
            "WILSONR = 2\nWILSONT = 2\n" % ((sun,) + dims + (beta, prec)))
    gens = (3, 8) if sun == 3 else (3, 2)
    last = _device_like_last(O, L, a, sun, dims, beta, gens=gens)
    r = _verify(tmp_path, init, a, last)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Verification successfully passed!" in r.stdout and "test failed" not in r.stdout
    for name in ("S_total", "Plq_total", "Polyakov_loop", "Polyakov_loop_P4", "Wilson_loop", "Fmunu_xy_3_re",
                 "Fmunu_abs_%d_re" % gens[1]):
        assert " GPU last %-20s:" % name in r.stdout and " CPU last %-20s:" % name in r.stdout
    assert " CPU l.varnc Fmunu_xy_3_re" in r.stdout
    bad = dict(last, Plq_total=last["Plq_total"] * (1 + 1e-4))
    r = _verify(tmp_path, init, a, bad)
    assert r.returncode == 4 and "Plq_total test failed!" in r.stdout and "Verification failed!" in r.stdout
    assert "S_total test failed" not in r.stdout
    near = dict(last, Polyakov_loop_P2=last["Polyakov_loop_P2"] * (1 + 1e-6))
    r = _verify(tmp_path, init, a, near)
    assert (r.returncode == 0) == (prec == "SINGLE"), r.stdout


def test_end_of_run_check_twisted_and_e_field(O, tmp_path):
    """With PHI / OMEGA the host check reads the twisted Y links at
    x = N1 - 1 as the device does (compared with the oracle's twisted view),
    and F0MU takes the (x,t), (y,t), (z,t) planes."""
    dims = (4, 4, 2, 4)
    L = O.lattice(3, dims, 5.9, 10)
    a = O.cold(L)
    rng = O.SeqRng("RANLUX", 8, 223)
    for _ in range(2):
        O.sweep(L, a, O.HB, rng)
    view = O.twisted_view(L, a, 0.7, -0.3)
    init = ("GROUP = 3\nL1 = 4\nL2 = 4\nL3 = 2\nL4 = 4\nBETA = 5.9\nPHI = 0.7\nOMEGA = -0.3\nF0MU = 1\n"
            "FMUNU1 = 1\nFMUNU6 = 1\nGETWILSON = 1\nWILSONR = 2\nWILSONT = 1\n")
    last = _device_like_last(O, L, view, 3, dims, 5.9, wilson=(2, 1), efield=True, gens=(1, 6))
    r = _verify(tmp_path, init, a, last)
    assert r.returncode == 0, r.stdout
    assert " GPU last Fmunu_zt_6_im" in r.stdout


@pytest.mark.gpu
def test_host_profile_table(tmp_path):
    """PROFILE = 1 prints GPU::print_time_detailed's kernel table
    (clinterface.cpp:1626-1666): one entry per kernel family with its launch
    count -- 8 heatbath and 8 NOR overrelaxation half-steps per sweep, one
    plaquette, Polyakov and final-sum launch per measurement -- and a mean
    time."""
    print("[case] host profile table", flush=True)
    f = tmp_path / "init.dat"
    f.write_text("GROUP = 3\nLS = 8\nLT = 4\nITER = 3\nNITER = 2\nNAV = 1\nBETA = 5.9\nINTS = 1\nPRNG = XOR128\n"
                 "PRECISION = SINGLE\nNOR = 1\nPL_LEVEL = 1\nTURNOFFFMUNU = 1\nPROFILE = 1\nOUTPUTPATH = %s/\n" % tmp_path)
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import re

    table = dict(re.findall(r'kernel "([^"]+)" executed (\d+) times', r.stdout))
    sweeps = 1 + 2 * 2   # NAV + (ITER - 1) NITER
    assert int(table["heatbath half-step"]) == 8 * sweeps, table
    assert int(table["overrelaxation half-step"]) == 8 * sweeps
    assert int(table["plaquette (lattice_measurement_plq)"]) == 3
    assert int(table["Polyakov loop (lattice_polyakov)"]) == 3
    assert int(table["final sums (reduce)"]) == 3
    times = [float(t) for t in re.findall(r"execution time: ([0-9.]+) \(\+/-", r.stdout)]
    assert len(times) == len(table) and all(t > 0 for t in times)
