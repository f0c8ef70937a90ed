"""bench.py's host-side pieces, on the CPU: the multi-GPU refusal, the
96-chain ensemble's composition, and the reference CPU timing harness that
the CPU baseline and the heatbath-only table run (oracle/_ref/ref_harness,
built from the reference sources where they are; skipped without it)."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402

HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "--gpus 2 is not supported" in r.stderr and r.stdout == ""


def test_c3_ensemble_composition():
    chains = bench._c3_chains()
    assert len(chains) == 96
    assert chains == bench._c3_chains()          # fixed by its seed
    for k, (dims, beta, seed) in enumerate(chains):
        assert dims[0] == dims[1] == dims[2] and dims[0] in (12, 16, 20, 24) and dims[3] in (4, 6, 8)
        assert beta == pytest.approx(5.5 + 0.7 * k / 95) and seed == 1000 + k
    assert chains[0][1] == 5.5 and chains[-1][1] == pytest.approx(6.2)
    # every (N_s, N_t) pair of the grid occurs
    assert len({c[0] for c in chains}) == 12


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference CPU build (oracle/_ref/ref_harness) absent")
def test_reference_harness_timing_and_warmup():
    # two timed sweeps of SU(2) 4^4: the count, a positive rate, a plaquette
    (sec, rate), = bench._ref_wait([bench._ref_bench(2, (4, 4, 4, 4), 2.3, 54, 2)])
    assert sec > 0 and rate == pytest.approx(4 * 256 * 2 / sec, rel=1e-3)
    out = subprocess.run([HARNESS, "bench", "2", "4", "4", "4", "4", "2.3", "10", "RANLUX3", "54", "2", "0"],
                         capture_output=True, text=True, check=True).stdout.split()
    plq_cold = float(out[3])
    # warm-up sweeps run before the timed ones: 1 warm + 1 timed == 2 timed
    out = subprocess.run([HARNESS, "bench", "2", "4", "4", "4", "4", "2.3", "10", "RANLUX3", "54", "1", "1"],
                         capture_output=True, text=True, check=True).stdout.split()
    assert out[0] == "1" and float(out[3]) == plq_cold
    # progress lines: one per half-step, on stderr only
    r = subprocess.run([HARNESS, "bench", "3", "4", "4", "2", "2", "5.9", "10", "RANLUX3", "7", "2", "0", "1"],
                       capture_output=True, text=True, check=True)
    assert len(r.stdout.split()) == 4
    lines = r.stderr.splitlines()
    assert len(lines) == 16 and lines[0] == "ref_harness: sweep 0 odd direction 0 done"
    assert lines[-1] == "ref_harness: sweep 1 even direction 3 done"
