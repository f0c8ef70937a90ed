"""Batched ensembles (C3 pattern): every chain of a batch must evolve exactly
as the same chain run alone, whatever the kernel shapes and the autotuner
chose; the C3 workload runs and gives per-chain observables."""
import numpy as np
import pytest

import qcdgpu_amd as Q

CHAINS = [((8, 8, 8, 4), 5.7, 11), ((6, 6, 6, 4), 6.0, 12), ((8, 8, 8, 6), 5.9, 13), ((6, 6, 6, 8), 5.5, 14)]


def _alone(dims, beta, seed, prng, prec, cycles, n_or):
    with Q.Lattice(sun=3, dims=dims, beta=beta, precision=prec, prng=prng, ranlux_luxury=3, randseries=seed,
                   max_measurements=16) as lat:
        lat.prng_init()
        lat.init("hot")
        for _ in range(cycles):
            lat.run_cycle(n_hb=1, n_or=n_or)
        lat.sync()
        return lat.export_links(), lat.raw_measurements()


@pytest.mark.gpu
@pytest.mark.parametrize("prng,prec", [("RANLUX", "fp32"), ("XOR128", "fp64")])
@pytest.mark.parametrize("tune", ["default", "autotune", "s1p1", "s3p2", "s1p2", "mixed", "per-bucket"])
def test_batch_equals_chains_alone(prng, prec, tune):
    """Every kernel shape and unit placement (by hand, or as the autotuner
    picks them) gives each chain exactly the chain run alone, with the
    buckets' half-steps merged into one launch per shape (the default; in
    "mixed" the two buckets use different shapes, so each half-step is two
    launches) or launched per bucket ("per-bucket")."""
    print("[case] batch %s %s tune=%s" % (prng, prec, tune), flush=True)
    cycles, n_or = 2, 2
    with Q.Batch(CHAINS, precision=prec, prng=prng, max_measurements=16) as b:
        assert b.info() == {"chains": 4, "buckets": 2}
        b.init("hot")
        if tune == "autotune":
            t = b.autotune(repeats=2)
            assert all(s in (1, 3) for d in t for s in d["shape"])
            assert all(p in (0, 1, 2) for d in t for p in d["placement"])
        elif tune == "mixed":
            b.set_tuning(0, shape=(3, 1), placement=(1, 2))
            b.set_tuning(1, shape=(1, 3), placement=(2, 0))
        elif tune == "per-bucket":
            b.set_merge(False)
        elif tune != "default":
            shape, place = int(tune[1]), int(tune[3])
            for k in range(2):
                b.set_tuning(k, shape=(shape, shape), placement=(place, 2 - place if k else place))
        b.capture_cycle(n_hb=1, n_or=n_or)
        b.launch_cycle(cycles)
        b.sync()
        for k, (dims, beta, seed) in enumerate(CHAINS):
            lat = b.chain(k)
            got = lat.export_links()
            plq, poly = lat.raw_measurements()
            ref, (rplq, rpoly) = _alone(dims, beta, seed, prng, prec, cycles, n_or)
            assert np.array_equal(got, ref), "chain %d links" % k
            assert np.array_equal(plq, rplq) and np.array_equal(poly, rpoly), "chain %d records" % k


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_c3_batched_ensemble():
    """C3: 96 SU(3) chains, (N_s, N_t) drawn from {12,16,20,24} x {4,6,8} with
    seed 54, beta_i = 5.5 + 0.7 i / 95, FP32, RANLUX3 per-chain streams,
    300 cycles of 1 heatbath + 4 overrelaxation sweeps with the plaquette and
    Polyakov loop every cycle; per-chain plaquette and Polyakov
    susceptibility."""
    print("[case] C3 96 chains", flush=True)
    rng = np.random.default_rng(54)
    ns = rng.choice([12, 16, 20, 24], size=96)
    nt = rng.choice([4, 6, 8], size=96)
    chains = [((int(ns[i]),) * 3 + (int(nt[i]),), 5.5 + 0.7 * i / 95, 1000 + i) for i in range(96)]
    with Q.Batch(chains, precision="fp32", prng="RANLUX", max_measurements=320) as b:
        assert b.info()["chains"] == 96
        b.init("hot")
        tuning = b.autotune(repeats=4)
        b.capture_cycle(n_hb=1, n_or=4)
        ms = b.time_cycle(300)
        b.sync()
        V = sum(int(np.prod(c[0])) for c in chains)
        rate = 300 * 5 * 4 * V / (ms * 1e-3)
        # algorithmic bytes per cycle: 4V link updates per sweep at 9 x 72 B
        # (+ 2 x 112 B of RANLUX state for the heatbath), plaquette and
        # Polyakov 5V x 72 B
        alg = 4 * V * ((9 * 72 + 2 * 112) + 4 * 9 * 72) + 5 * V * 72
        gbs = alg * 300 / (ms * 1e-3) / 1e9
        print("  tuning %s" % tuning)
        print("  %.1f ms for 300 cycles, %.3e link updates/s, %.0f GB/s algorithmic = %.1f %% of 8 TB/s"
              % (ms, rate, gbs, gbs / 80.0))
        plaq, chi = [], []
        for k, (dims, beta, _) in enumerate(chains):
            m = b.chain(k).measurements()
            assert len(m["plq_total"]) == 300
            plaq.append(np.mean(m["plq_total"][100:]) / 3)
            chi.append(Q.polyakov_susceptibility({x: m[x][100:] for x in ("polyakov_re", "polyakov_im")}, dims, 3))
        plaq = np.array(plaq)
        betas = np.array([c[1] for c in chains])
        for k in (0, 47, 95):
            print("  chain %2d %s beta %.3f  <plaq>/3 %.5f  chi_P %.4f" % (k, chains[k][0], betas[k], plaq[k], chi[k]))
        # the plaquette grows with beta across the whole ensemble
        assert np.corrcoef(betas, plaq)[0, 1] > 0.9
        assert np.all(np.isfinite(chi)) and np.all(np.array(chi) >= 0)
