"""The oracle's C4 plaquette fixture (tests/golden/c4_oracle_plaquette.json,
written by tests/golden/make_c4_oracle_plaquette.py): eight independent FP64
chains whose pooled mean is what the GPU's plain heatbath ensemble is checked
against.  Checked here: the pooled figures follow from the chains, the
chains scatter as their own binned errors say (chi-square of the chain
means about the pooled mean), and the value lies between the first-order
strong- and weak-coupling expansions of the SU(2) plaquette at beta = 2.2,
as the measured plaquette there does."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_c4_oracle_fixture_consistent():
    with open(os.path.join(HERE, "golden", "c4_oracle_plaquette.json")) as f:
        d = json.load(f)
    assert d["lattice"] == [8, 8, 8, 4] and d["beta"] == 2.2 and d["group"] == 2
    means = np.array([c["mean"] for c in d["chains"]])
    errs = np.array([c["err"] for c in d["chains"]])
    assert len(means) == 8 and len({c["randseries"] for c in d["chains"]}) == 8
    assert abs(d["pooled_mean"] - means.mean()) < 1e-15
    assert abs(d["pooled_err"] - np.sqrt((errs ** 2).sum()) / 8) < 1e-15
    chi2 = (((means - means.mean()) / errs) ** 2).sum()
    assert 0.5 < chi2 < 25.0   # 7 degrees of freedom: p in (1e-4, 1 - 1e-4)
    # first order in beta: beta/4 = 0.55; first order in 1/beta:
    # 1 - 3/(4 beta) = 0.659
    assert 0.55 < d["pooled_mean"] < 0.659
    assert all(3.0 < c["tau_int"] < 6.0 for c in d["chains"])
