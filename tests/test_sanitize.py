"""Memory and undefined-behaviour checks of the CPU code, built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++ / gcc, host code only,
-fno-sanitize-recover so the first report fails the run):

- the oracle, through oracle/sanitize_main.c (every family of its routines
  on lattices whose four extents all differ);
- the command-line host program: the init.dat parser (the reference's own
  init.dat and a file with every key family), the end-of-run check of a
  configuration (--verify, SU(3) and SU(2), the F_munu and Wilson columns),
  and a .qcg header read (INIT = 0).
No GPU is used; the host program loads libqcdhip.so but calls no device
function on these paths."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "qcdgpu_amd")
ORACLE = os.path.join(REPO, "oracle")
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_host import _device_like_last  # noqa: E402


def _have(tool):
    return shutil.which(tool) is not None


@pytest.fixture(scope="module")
def host_asan(tmp_path_factory):
    if not os.path.exists(os.path.join(PKG, "libqcdhip.so")):
        pytest.skip("libqcdhip.so not built")
    out = str(tmp_path_factory.mktemp("asan") / "qcdgpu_hip_asan")
    cmd = (["g++", "-std=c++17"] + SAN + ["-o", out]
           + [os.path.join(PKG, "host", n) for n in ("qcdgpu_main.cpp", "host_check.cpp")]
           + ["-L" + PKG, "-lqcdhip", "-Wl,-rpath," + PKG, "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _clean(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text, text[-4000:]
    return text


@pytest.mark.skipif(not _have("gcc"), reason="gcc missing")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    r = subprocess.run(["gcc", "-std=c11"] + SAN + ["-I" + ORACLE, os.path.join(ORACLE, "sanitize_main.c"),
                        os.path.join(ORACLE, "oracle.c"), "-lm", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    _clean(r)
    assert r.returncode == 0
    assert len([x for x in r.stdout.splitlines() if x.startswith("SU(")]) == 6


def test_host_parser_under_sanitizers(host_asan, tmp_path):
    ref = os.path.join(REPO, "tests", "golden", "reference_init.dat")
    r = subprocess.run([host_asan, "--print-params", ref], capture_output=True, text=True, env=ENV)
    _clean(r)
    assert r.returncode in (0, 2) and "GROUP=" in r.stdout
    f = tmp_path / "init.dat"
    f.write_text("\tGROUP = 2\n\tLT = 8\n\tLS = 12\n\tL1 = 16   # x\n\tBETA = 2.3e+00\n\tNAV = 5\n\tINTS = 0\n"
                 "PRNG = RANLUX2\nPRECISION = SINGLE\nNOR = 3\nPHI = 0.5\nOMEGA = -0.25\nFMUNU = 1\nFMUNU3 = 1\n"
                 "FMUNU8 = 1\nGETWILSON = 1\nWILSONR = 3\nWILSONT = 2\nPL = 3\nACTIONS_DIFF = 1\n"
                 "OUTPUTPATH = " + str(tmp_path) + "/\nPRNG_INSTANCES = 128\nPROFILE = 1\nUNKNOWNKEY = 7\n")
    r = subprocess.run([host_asan, "--print-params", str(f)], capture_output=True, text=True, env=ENV)
    _clean(r)
    assert "GROUP=2" in r.stdout


@pytest.mark.parametrize("sun,beta", [(3, 5.9), (2, 2.3)])
def test_host_verify_under_sanitizers(O, host_asan, tmp_path, sun, beta):
    dims = (4, 6, 2, 4)
    L = O.lattice(sun, dims, beta, 10)
    a = O.cold(L)
    rng = O.SeqRng("RANLUX", 5, 223)
    for _ in range(2):
        O.sweep(L, a, O.HB, rng)
    init = ("GROUP = %d\nL1 = %d\nL2 = %d\nL3 = %d\nL4 = %d\nBETA = %r\nPRECISION = DOUBLE\nGETWILSON = 1\n"
            "WILSONR = 2\nWILSONT = 2\n" % ((sun,) + dims + (beta,)))
    last = _device_like_last(O, L, a, sun, dims, beta, gens=(3, 8) if sun == 3 else (3, 2))
    (tmp_path / "init.dat").write_text(init)
    np.ascontiguousarray(a, dtype=np.float64).tofile(tmp_path / "links.f64")
    (tmp_path / "last.txt").write_text("".join("%s %.17g\n" % kv for kv in last.items()))
    r = subprocess.run([host_asan, "--verify", str(tmp_path / "init.dat"), str(tmp_path / "links.f64"),
                        str(tmp_path / "last.txt")], capture_output=True, text=True, env=ENV)
    _clean(r)
    assert r.returncode == 0 and "Verification successfully passed!" in r.stdout
    # a short links file and a missing values file are errors, not reads past the end
    (tmp_path / "short.f64").write_bytes(b"\0" * 64)
    r = subprocess.run([host_asan, "--verify", str(tmp_path / "init.dat"), str(tmp_path / "short.f64"),
                        str(tmp_path / "last.txt")], capture_output=True, text=True, env=ENV)
    _clean(r)
    assert r.returncode == 1
    r = subprocess.run([host_asan, "--verify", str(tmp_path / "init.dat"), str(tmp_path / "links.f64"),
                        str(tmp_path / "missing.txt")], capture_output=True, text=True, env=ENV)
    _clean(r)
    assert r.returncode == 1


def test_host_qcg_header_under_sanitizers(host_asan, tmp_path):
    import qcg

    rng = np.random.default_rng(1)
    dims = (4, 4, 4, 2)
    z = rng.normal(size=(128, 4, 3, 3)) + 1j * rng.normal(size=(128, 4, 3, 3))
    q, _ = np.linalg.qr(z)
    links = np.stack([q.real, q.imag], axis=-1)
    h = qcg.make_header(3, dims, 5.9, "single", "XOR128", iter=4, iter_counter=2)
    (tmp_path / "s.qcg").write_bytes(qcg.encode(h, links))
    (tmp_path / "init.dat").write_text("INIT = 0\nSTATEFILE = %s\nITER = 4\n" % (tmp_path / "s.qcg"))
    r = subprocess.run([host_asan, "--print-params", str(tmp_path / "init.dat")], capture_output=True, text=True,
                       env=ENV)
    _clean(r)
    assert r.returncode == 0 and "GROUP=3" in r.stdout and "BETA=5.9" in r.stdout
    (tmp_path / "t.qcg").write_bytes(qcg.encode(h, links)[:200])
    (tmp_path / "init2.dat").write_text("INIT = 0\nSTATEFILE = %s\n" % (tmp_path / "t.qcg"))
    r = subprocess.run([host_asan, "--print-params", str(tmp_path / "init2.dat")], capture_output=True, text=True,
                       env=ENV)
    _clean(r)
    assert r.returncode == 2
