"""Writes tests/golden/c4_oracle_plaquette.json: the mean plaquette of the
C4 lattice (SU(2) 8^3 x 4, beta 2.2) from the oracle -- FP64 links, the
reference CPU path's semantics (cold start, sequential RANLUX3 numbers,
Kennedy-Pendleton heatbath, suncpp's sweep order) -- in 8 independent
chains (random series 101..108) of 300 thermalisation and 20000 measured
sweeps each, per chain the mean of ReTr U_p / 2 over the measured sweeps
and its binned error (20 bins).  About four minutes on 8 cores:

    python tests/golden/make_c4_oracle_plaquette.py
"""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)

DIMS, BETA, THERM, SWEEPS = (8, 8, 8, 4), 2.2, 300, 20000


def chain(seed):
    import pyoracle as O
    from qcdgpu_amd import analysis as A

    L = O.lattice(2, DIMS, BETA, 10)
    a = O.cold(L)
    rng = O.SeqRng("RANLUX", seed, 223)
    V = int(np.prod(DIMS))
    for _ in range(THERM):
        O.sweep(L, a, O.HB, rng)
    p = np.empty(SWEEPS)
    for k in range(SWEEPS):
        O.sweep(L, a, O.HB, rng)
        p[k] = O.plaquette(L, a).sum() / (6 * V) / 2.0
    m, e = A.binned_error(p, nbins=20)
    return {"randseries": seed, "mean": m, "err": e, "tau_int": A.tau_int(p)[0]}


def main():
    with Pool(8) as pool:
        chains = pool.map(chain, range(101, 109))
    means = np.array([c["mean"] for c in chains])
    errs = np.array([c["err"] for c in chains])
    out = {"lattice": list(DIMS), "beta": BETA, "group": 2, "thermalisation": THERM, "sweeps": SWEEPS,
           "observable": "mean over sweeps of sum ReTr U_p / (6 V) / 2",
           "chains": chains, "pooled_mean": means.mean(), "pooled_err": float(np.sqrt((errs ** 2).sum()) / len(errs))}
    with open(os.path.join(HERE, "c4_oracle_plaquette.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["pooled_mean"], out["pooled_err"])


if __name__ == "__main__":
    main()
