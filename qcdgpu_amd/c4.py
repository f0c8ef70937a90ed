"""C4 workload: learned proposal vs plain heatbath on SU(2) 8^3 x 4, beta 2.2.

Three steps, each its own process (the training step runs PyTorch on the
GPU, the others libqcdhip.so; no process loads both):

  python -m qcdgpu_amd.c4 gen   --out data.npy    heatbath configurations
  python -m qcdgpu_amd.c4 train --data data.npy --out params.npy
  python -m qcdgpu_amd.c4 chain --params params.npy   plain vs learned chains

`gen` runs the plain heatbath of the C0 form (Haar hot start, 1 Kennedy-
Pendleton + 3 overrelaxation sweeps per configuration) and keeps every
configuration after thermalisation.  `train` fits the flow by forward KL.
`chain` runs chains from the same start for the same number of steps:
plain = 1 heatbath sweep per step; learned = 1 heatbath sweep and one
learned Metropolis sweep per step (every coupling layer proposes its active
links from its map of Haar draws, each accepted with the exact
Wilson-action Metropolis probability, all on the device); with --global
also the whole-lattice independence-Metropolis step with the flow.  Every
arm measures the plaquette, the Polyakov loop and the 2x2 Wilson loop every
step.  It prints one JSON line: observables with binned errors and their
difference from the plain arm in standard deviations, the Polyakov loop's
tau_int, steps per second and independent configurations per second of
each arm, and the acceptances.
"""
import argparse
import json
import sys
import time

import numpy as np

DIMS = (8, 8, 8, 4)
BETA = 2.2


def cmd_gen(a):
    import qcdgpu_amd as Q
    from qcdgpu_amd import flow as QF

    dims = tuple(a.dims)
    out = np.lib.format.open_memmap(a.out, mode="w+", dtype=np.float32, shape=(a.n, 4, int(np.prod(dims)), 4))
    t0 = time.perf_counter()
    with Q.Lattice(sun=2, dims=dims, beta=a.beta, precision="fp32", prng="XOR128", randseries=a.seed,
                   max_measurements=4) as lat:
        lat.prng_init()
        lat.init("hot")
        for _ in range(a.therm):
            lat.sweep("hb")
            for _ in range(3):
                lat.sweep("or")
        for i in range(a.n):
            lat.sweep("hb")
            for _ in range(3):
                lat.sweep("or")
            out[i] = QF.to_quaternions(lat.export_links(), dims)
            if (i + 1) % 5000 == 0:
                print("  gen %d of %d (%.1f s)" % (i + 1, a.n, time.perf_counter() - t0), flush=True)
    out.flush()
    print(json.dumps({"gen_configs": a.n, "gen_seconds": round(time.perf_counter() - t0, 2)}), flush=True)


def cmd_train_device(a):
    """Layer-local forward KL with the device's training kernels."""
    from qcdgpu_amd import flow as QF

    dims = tuple(a.dims)
    data = np.load(a.data, mmap_mode="r")
    X = np.array(data[: a.n])
    rng = np.random.default_rng(a.seed)
    with QF.Flow(dims=dims, layers=a.layers, bins=a.bins, max_batch=a.batch) as f:
        f.set_params(np.zeros(f.param_count(), dtype=np.float32))
        t0 = time.perf_counter()
        hist = []
        for it in range(a.steps):
            idx = rng.choice(len(X), size=a.batch, replace=False)
            hist.append(f.train_step(X[idx], a.lr))
            if it % max(a.steps // 10, 1) == 0 or it == a.steps - 1:
                print("  step %5d  layer-local -log q %.3f" % (it, hist[-1]), flush=True)
        secs = time.perf_counter() - t0
        np.save(a.out, f.get_params())
    print(json.dumps({"train_configs": int(len(X)), "train_steps": a.steps, "batch": a.batch, "trainer": "device",
                      "train_seconds": round(secs, 1), "loss_first": round(hist[0], 3),
                      "loss_last": round(float(np.mean(hist[-10:])), 3)}), flush=True)


def cmd_train(a):
    if a.lr is None:
        a.lr = 1e-4 if a.torch else 1e-3
    if not a.torch:
        return cmd_train_device(a)
    import torch

    from qcdgpu_amd import flow as QF

    dev = "cuda" if torch.cuda.is_available() else "cpu"
    data = np.load(a.data, mmap_mode="r")
    X = torch.as_tensor(np.array(data[: a.n]), device=dev)
    m = QF.FlowModel(dims=tuple(a.dims), layers=a.layers, bins=a.bins, device=dev, seed=a.seed)
    with torch.no_grad():
        base = float(-m.log_q(X[:64]).mean())
    t0 = time.perf_counter()
    hist = QF.train_forward_kl(m, X, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log_every=max(a.steps // 10, 1))
    secs = time.perf_counter() - t0
    with torch.no_grad():
        held = float(-m.log_q(X[-64:]).mean())
    np.save(a.out, m.flat_params())
    print(json.dumps({"train_configs": int(X.shape[0]), "train_steps": a.steps, "batch": a.batch,
                      "train_seconds": round(secs, 1), "neg_log_q_identity": round(base, 3),
                      "neg_log_q_trained": round(held, 3), "last_loss": round(hist[-1], 3)}), flush=True)


def _stats(series, sweeps_per_s):
    from qcdgpu_amd import analysis as A

    m, e = A.binned_error(series, nbins=20)
    tau = A.tau_int(series)[0]
    return m, e, tau, sweeps_per_s / (2.0 * tau)


def run_arm(arm, a, flow, seed=None):
    """One chain: "plain" (heatbath), "learned" (heatbath + learned
    Metropolis sweep) or "global" (heatbath + whole-lattice independence
    Metropolis).  Returns the arm's result dict."""
    import qcdgpu_amd as Q
    from qcdgpu_amd import flow as QF

    dims = tuple(a.dims)
    if seed is None:
        seed = a.seed + {"plain": 0, "learned": 1000, "global": 2000}[arm]
    with Q.Lattice(sun=2, dims=dims, beta=a.beta, precision="fp32", prng="XOR128", randseries=seed,
                   max_measurements=a.steps + 2) as lat:
        lat.set_wilson(2, 2)
        lat.prng_init()
        lat.init("hot")
        for _ in range(a.therm):
            lat.sweep("hb")
        lat.reset_measurements()
        lw = {}
        if arm in ("plain", "learned"):
            # one step (1 heatbath sweep [+ 1 learned sweep] + measurements)
            # recorded once as a graph and replayed; the learned sweeps take
            # successive seeds (seed << 32) + 1, + 2, ... from the device
            if arm == "learned":
                flow.imh_seed(seed << 32)
                flow.imh_counts(reset=True)
            lat.graph_begin()
            lat.sweep("hb")
            if arm == "learned":
                flow.imh_sweep_next(lat, a.beta)
            lat.measure(wilson=True)
            lat.graph_end()
            lat.sync()
            t0 = time.perf_counter()
            lat.launch_cycle(a.steps)
            lat.sync()
            secs = time.perf_counter() - t0
            acc = 0.0
            if arm == "learned":
                acc_n, prop_n = flow.imh_counts()
                acc = acc_n / max(prop_n, 1)
        else:
            acc, secs = QF.imh_chain(lat, flow if arm == "global" else None, a.beta, a.steps, hb_per_step=1,
                                     seed=seed + 17, wilson=True, log_weights=lw)
        m = lat.measurements()
        plq = m["plq_total"] / 2.0
        pol = m["polyakov_re"]
        wil = lat.wilson_measurements() / 2.0
        rate = a.steps / secs
        out = {"steps_per_s": round(rate, 1), "acceptance": round(acc, 4)}
        for name, s in (("plaquette", plq), ("polyakov", pol), ("wilson_2x2", wil)):
            mu, err, tau, ind = _stats(s, rate)
            out[name] = {"mean": mu, "err": err, "tau_int": round(tau, 2)}
            if name == "polyakov":
                out["independent_configs_per_s"] = round(ind, 2)
        out["tau_int_abs_polyakov"] = round(_stats(np.abs(pol), rate)[2], 2)
        for k, v in lw.items():   # log w = -S - log q: the acceptance compares proposal with current
            v = np.asarray(v)
            out["log_w_" + k] = {"mean": round(float(v.mean()), 2), "std": round(float(v.std()), 2)}
        return out


def cmd_chain(a):
    from qcdgpu_amd import flow as QF

    dims = tuple(a.dims)
    res = {}
    flow = None
    if a.params:
        flow = QF.Flow(dims=dims, layers=a.layers, bins=a.bins, max_batch=a.batch)
        flow.set_params(np.load(a.params).astype(np.float32))
    arms = ["plain"] + (["learned"] + (["global"] if a.global_imh else []) if flow is not None else [])
    # The observables are checked against the plain heatbath ensemble: chains
    # with random series 1 .. plain_chains, pooled (one chain can sit 3
    # sigma off; the pool's error is what a comparison at 2 sigma needs).
    # Its rate and independent configurations per second are the chains'
    # means.
    plain = []
    for k in range(1, a.plain_chains + 1):
        plain.append(run_arm("plain", a, None, seed=k))
        print("  plain %d: %s" % (k, json.dumps(plain[-1])), flush=True)
    pooled = {}
    for name in ("plaquette", "polyakov", "wilson_2x2"):
        ms = np.array([c[name]["mean"] for c in plain])
        es = np.array([c[name]["err"] for c in plain])
        pooled[name] = {"mean": float(ms.mean()), "err": float(np.sqrt((es ** 2).sum()) / len(es)),
                        "tau_int": round(float(np.mean([c[name]["tau_int"] for c in plain])), 2)}
    res["plain"] = dict(pooled, chains=len(plain),
                        steps_per_s=round(float(np.mean([c["steps_per_s"] for c in plain])), 1),
                        independent_configs_per_s=round(float(np.mean([c["independent_configs_per_s"]
                                                                        for c in plain])), 2))
    for arm in arms[1:]:
        res[arm] = run_arm(arm, a, flow)
        print("  %s: %s" % (arm, json.dumps(res[arm])), flush=True)
    z = {}
    for arm in arms[1:]:
        z[arm] = {}
        for name in ("plaquette", "polyakov", "wilson_2x2"):
            p, l = res["plain"][name], res[arm][name]
            z[arm][name] = round((l["mean"] - p["mean"]) / np.hypot(p["err"], l["err"]), 2)
    best = "learned" if "learned" in res else "plain"
    line = {
        "metric": "independent configurations/s (Polyakov loop tau_int), learned + heatbath vs plain heatbath",
        "value": res[best]["independent_configs_per_s"],
        "unit": "configurations/s",
        "plain": res["plain"]["independent_configs_per_s"],
        "ratio": round(res[best]["independent_configs_per_s"] / res["plain"]["independent_configs_per_s"], 3),
        "acceptance": {arm: res[arm]["acceptance"] for arm in arms[1:]},
        "z_vs_plain": z,
        "config": {"workload": "SU(2) %dx%dx%dx%d beta=%.2f, FP32 links" % (dims + (a.beta,)), "steps": a.steps,
                   "thermalisation": a.therm, "flow": {"layers": a.layers, "bins": a.bins, "hidden": 32}},
        "arms": res,
    }
    print(json.dumps(line), flush=True)
    if flow is not None:
        flow.close()


def main(argv=None):
    p = argparse.ArgumentParser(prog="python -m qcdgpu_amd.c4")
    sub = p.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--out", required=True)
    g.add_argument("--n", type=int, default=20000)
    g.add_argument("--therm", type=int, default=200)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--n", type=int, default=20000)
    t.add_argument("--steps", type=int, default=1000)
    t.add_argument("--batch", type=int, default=16)
    # each output of the second convolution sums 2592 inputs: Adam steps of
    # 1e-3 per weight move the spline parameters by O(1)
    t.add_argument("--lr", type=float, default=None, help="Adam step (default 1e-3 device, 1e-4 PyTorch)")
    t.add_argument("--torch", action="store_true",
                   help="train the whole-flow forward KL with the PyTorch twin instead of the device kernels")
    c = sub.add_parser("chain")
    c.add_argument("--params", default=None)
    c.add_argument("--steps", type=int, default=4000)
    c.add_argument("--therm", type=int, default=200)
    c.add_argument("--batch", type=int, default=64)
    c.add_argument("--global", dest="global_imh", action="store_true",
                   help="also run the whole-lattice independence-Metropolis arm")
    c.add_argument("--plain-chains", type=int, default=8,
                   help="plain heatbath chains (random series 1..n) pooled as the reference ensemble")
    for q in (g, t, c):
        q.add_argument("--dims", type=int, nargs=4, default=list(DIMS))
        q.add_argument("--beta", type=float, default=BETA)
        q.add_argument("--seed", type=int, default=54)
        q.add_argument("--layers", type=int, default=8)
        q.add_argument("--bins", type=int, default=4)
    a = p.parse_args(argv)
    {"gen": cmd_gen, "train": cmd_train, "chain": cmd_chain}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
