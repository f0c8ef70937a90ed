"""Headline benchmark: SU(3) link updates per second on one MI355X.

Workload (one "step" = one sweep of the measurement cycle):
  SU(3) Wilson action, 32^3 x 8 lattice, beta = 5.9, FP32 links stored as
  18 reals (72 B), FP64 reunitarisation and FP64 measurement sums, cold start,
  XOR128 per-site streams in reference stream layout, NHIT = 10.
  One sweep = 1 Cabibbo-Marinari heatbath sweep + 4 overrelaxation sweeps +
  plaquette + Polyakov loop, captured once as a hipGraph and replayed.
  Link updates per sweep = 5 * 4 * V (V = 262144).

Prints one JSON line.  `roofline` is for the dominant kernel (the
overrelaxation half-step, 32 of the 40 update launches of a sweep):
algorithmic bytes per launch = 9 * 72 B * V/2 (each updated link reads the
three other directions' links of both parities and its own direction's
links of both parities once, and writes itself), divided by the kernel's
duration: HIP events on the lattice stream around each kernel of eagerly
launched cycles, scaled so that a cycle's kernels add up to the graph-timed
step.  `config.step_roofline` is the whole step's algorithmic bytes over the
timed step.  `cpu_baseline` times the reference CPU path (suncpp, built from
the reference sources into oracle/_ref/) on the host's cores for a bounded
sample of heatbath sweeps of the same lattice.

Other lines: --c3 (the 96-chain batched ensemble), --c4 (the learned
proposal), --table (heatbath-only CPU and GPU rows of SU(2) 8^4, SU(3)
32^3x8, SU(3) 64^4 and the 96-chain ensemble), --copy-calibration.
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

DIMS = (32, 32, 32, 8)
BETA = 5.9
N_OR = 4
LINK_BYTES = 72
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "or_halfstep_su3_32c8_fp32_traffic.json")


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _harness(name, args, timeout=300):
    out = subprocess.run([os.path.join(REPO, "oracle", "_ref", name)] + args, capture_output=True, text=True,
                         timeout=timeout, check=True).stdout.split()
    return float(out[2])   # link updates per second


def cpu_baseline(sweeps=4, o3_sweeps=2, agg_sweeps=2):
    """The reference's own CPU path (suncpp, single-threaded FP64, RANLUX3,
    heatbath only; compiled from the reference sources into oracle/_ref/):
    one core at the reference's -O2 and at -O3 (x86-64-v3), the all-core
    aggregate of independent processes with distinct seeds (suncpp has no
    threads), and, labelled as an extension, the oracle's overrelaxation
    sweep (suncpp has none).  Without the reference build the baseline is
    the oracle port, and the line says why."""
    V = 1
    for d in DIMS:
        V *= d
    lat = [str(d) for d in DIMS]
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))   # one process per core, at most the box's CPU share
    info = {"nproc": os.cpu_count(), "cpu_model": _cpu_model(), "cores_available": cores}
    sample = "%d cold-start Cabibbo-Marinari heatbath sweeps of SU(3) %dx%dx%dx%d beta=%.1f, FP64, RANLUX3 (suncpp)" % (
        (sweeps,) + DIMS + (BETA,)
    )
    harness = os.path.join(REPO, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        print("bench.py: WARNING: the reference CPU build oracle/_ref/ref_harness is missing (built by "
              "oracle/build_ref.sh where the reference sources are); cpu_baseline falls back to the oracle port",
              file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle as O

        L = O.lattice(3, DIMS, BETA, 10)
        a = O.cold(L)
        rng = O.SeqRng("RANLUX", 54, 223)
        t0 = time.perf_counter()
        O.sweep(L, a, O.HB, rng)
        dt = time.perf_counter() - t0
        return dict(info, value=4.0 * V / dt, unit="link updates/s", cores=1, kind="port",
                    fallback_reason="oracle/_ref/ref_harness missing",
                    sample="1 cold-start heatbath sweep of SU(3) 32x32x32x8, FP64 oracle restatement of suncpp")
    base = ["bench", "3"] + lat + [str(BETA), "10", "RANLUX3"]
    o2 = _harness("ref_harness", base + ["54", str(sweeps)])
    rows = {"O2_1core": o2}
    o3_path = os.path.join(REPO, "oracle", "_ref", "ref_harness_o3")
    if os.path.exists(o3_path):
        rows["O3_x86-64-v3_1core"] = _harness("ref_harness_o3", base + ["54", str(o3_sweeps)])
    # all-core aggregate: one process per core, distinct random series
    procs = [subprocess.Popen([harness] + base + [str(54 + k), str(agg_sweeps)], stdout=subprocess.PIPE, text=True)
             for k in range(cores)]
    agg = 0.0
    for pr in procs:
        out, _ = pr.communicate(timeout=600)
        if pr.returncode != 0:
            raise RuntimeError("ref_harness failed in the all-core run")
        agg += float(out.split()[2])
    rows["O2_all_cores"] = agg
    # extension: overrelaxation (not in suncpp), the oracle's FP64 C restatement
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as O

    L = O.lattice(3, DIMS, BETA, 10)
    a = O.cold(L)
    t0 = time.perf_counter()
    O.sweep(L, a, O.OR, O.GidRng())
    rows["extension_overrelaxation_port_1core"] = 4.0 * V / (time.perf_counter() - t0)
    return dict(info, value=o2, unit="link updates/s", cores=1, kind="reference", sample=sample,
                rows=rows, aggregate_cores=cores,
                rows_sample="O3: %d sweeps; all cores: %d sweeps per process, random series 54 + k; "
                            "overrelaxation: 1 sweep of the oracle (FP64 C)" % (o3_sweeps, agg_sweeps))


def _ref_bench(sun, dims, beta, seed, sweeps, warm=0, progress=False):
    """One ref_harness bench process timing `sweeps` heatbath sweeps after
    `warm` untimed ones (with `progress`, a stderr line per half-step)."""
    exe = os.path.join(REPO, "oracle", "_ref", "ref_harness")
    return subprocess.Popen([exe, "bench", str(sun)] + [str(d) for d in dims]
                            + [repr(float(beta)), "10", "RANLUX3", str(seed), str(sweeps), str(warm),
                               "1" if progress else "0"],
                            stdout=subprocess.PIPE, text=True)


def _ref_wait(procs, timeout=900):
    """(seconds, link updates/s) of each process, in order."""
    out = []
    for pr in procs:
        text, _ = pr.communicate(timeout=timeout)
        if pr.returncode != 0:
            raise RuntimeError("ref_harness failed (exit %d)" % pr.returncode)
        f = text.split()
        out.append((float(f[1]), float(f[2])))
    return out


def _c3_chains():
    import numpy as np

    rng = np.random.default_rng(54)
    ns = rng.choice([12, 16, 20, 24], size=96)
    nt = rng.choice([4, 6, 8], size=96)
    return [((int(ns[i]),) * 3 + (int(nt[i]),), 5.5 + 0.7 * i / 95, 1000 + i) for i in range(96)]


def baseline_table(args):
    """Four configurations (SU(2) 8^4 beta 2.3; SU(3) 32^3x8 beta 5.9;
    SU(3) 64^4 beta 6.0; the 96-chain SU(3) ensemble of --c3), each as
    heatbath-only sweeps with a plaquette record: the
    reference's CPU path (suncpp -O2, FP64, RANLUX3) on one core and as the
    aggregate of one process per core, and this library on the GPU at FP32
    and FP64.  Prints one JSON line per configuration and GPU precision.
    Heatbath-only on both sides because suncpp has no overrelaxation; the
    headline line covers the 1 heatbath + 4 overrelaxation workload."""
    import numpy as np

    import qcdgpu_amd as Q

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    host = {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "processes_all_core": cores}
    singles = [
        ("C0", 2, (8, 8, 8, 8), 2.3, 2000, 200, 200, 20, cores),
        ("C1", 3, DIMS, BETA, 400, 40, 20, 2, cores),
        # 64^4 FP64 is 9.7 GB of links per suncpp process: the aggregate uses
        # at most 8 processes so the host memory stays bounded.
        ("C2", 3, (64, 64, 64, 64), 6.0, 40, 4, 1, 0, min(cores, 8)),
    ]
    have_ref = os.path.exists(os.path.join(REPO, "oracle", "_ref", "ref_harness"))
    if not have_ref:
        sys.exit("bench.py --table: oracle/_ref/ref_harness is missing (oracle/build_ref.sh builds it where the "
                 "reference sources are)")

    def hb_bytes(V, s, state):
        # a heatbath sweep updates 4V links at 9 s (+ the stream state read
        # and written) and the plaquette reads the 4V links once
        return 4.0 * V * (9 * s + 2 * state) + 4.0 * V * s

    for name, sun, dims, beta, steps, warm, cpu_sweeps, cpu_warm, nproc in singles:
        V = int(np.prod(dims))
        for prec in ("fp32", "fp64"):
            print("bench.py --table: %s GPU %s" % (name, prec), file=sys.stderr, flush=True)
            with Q.Lattice(sun=sun, dims=dims, beta=beta, precision=prec, prng="XOR128", randseries=54,
                           max_measurements=steps + warm + 8) as lat:
                lat.init("cold")
                lat.prng_init()
                lat.capture_cycle(n_hb=1, n_or=0, plaquette=True, polyakov=False)
                lat.launch_cycle(warm)
                lat.sync()
                ms = lat.time_cycle(steps)
                plq = lat.measurements()["plq_total"][-1] / sun
                s = (lat.params.storage_reals) * (4 if prec == "fp32" else 8)
            sweep_s = steps / (ms * 1e-3)
            gbs = hb_bytes(V, s, 16) * sweep_s / 1e9
            row = {"config": name, "group": "SU(%d)" % sun, "lattice": list(dims), "beta": beta,
                   "gpu": {"precision": prec, "prng": "XOR128", "sweeps_per_s": sweep_s,
                           "link_updates_per_s": 4.0 * V * sweep_s, "hbm_frac": gbs / HBM_PEAK_GBS,
                           "achieved_GBs": gbs, "algorithmic_bytes_per_sweep": hb_bytes(V, s, 16),
                           "sweeps_timed": steps, "plaquette_last": float(plq)}}
            if prec == "fp64":
                print("bench.py --table: %s CPU 1 core, %d sweeps" % (name, cpu_sweeps), file=sys.stderr, flush=True)
                big = V > 1 << 22   # a 64^4 sweep runs for minutes: progress lines
                (sec, rate), = _ref_wait([_ref_bench(sun, dims, beta, 54, cpu_sweeps, cpu_warm, progress=big)])
                print("bench.py --table: %s CPU %d processes" % (name, nproc), file=sys.stderr, flush=True)
                t0 = time.perf_counter()
                agg = _ref_wait([_ref_bench(sun, dims, beta, 54 + k, cpu_sweeps, cpu_warm, progress=big and k == 0)
                                 for k in range(nproc)])
                wall = time.perf_counter() - t0
                row["cpu"] = dict(host, kind="reference", build="suncpp -g -O2, FP64, RANLUX3",
                                  sweeps_timed=cpu_sweeps, warmup=cpu_warm,
                                  one_core_sweeps_per_s=cpu_sweeps / sec, one_core_link_updates_per_s=rate,
                                  all_core_processes=nproc, all_core_link_updates_per_s=sum(r for _, r in agg),
                                  all_core_wall_s=wall)
            print(json.dumps(row), flush=True)

    chains = _c3_chains()
    Vs = [int(np.prod(c[0])) for c in chains]
    V = sum(Vs)
    for prec in ("fp32", "fp64"):
        print("bench.py --table: C3 GPU %s" % prec, file=sys.stderr, flush=True)
        with Q.Batch(chains, precision=prec, prng="RANLUX", max_measurements=300) as b:
            b.init("hot")
            b.autotune(repeats=4)
            b.set_merge(True)
            b.capture_cycle(n_hb=1, n_or=0, plaquette=True, polyakov=False)
            b.launch_cycle(20)
            b.sync()
            ms = b.time_cycle(200)
            b.sync()
        s = 72 if prec == "fp32" else 144
        sweep_s = 200 / (ms * 1e-3)
        gbs = hb_bytes(V, s, 112) * sweep_s / 1e9
        row = {"config": "C3", "group": "SU(3)", "chains": 96, "sites": V, "beta": [5.5, 6.2],
               "gpu": {"precision": prec, "prng": "RANLUX3", "ensemble_sweeps_per_s": sweep_s,
                       "link_updates_per_s": 4.0 * V * sweep_s, "hbm_frac": gbs / HBM_PEAK_GBS,
                       "achieved_GBs": gbs, "algorithmic_bytes_per_sweep": hb_bytes(V, s, 112),
                       "sweeps_timed": 200}}
        if prec == "fp64":
            # 10 sweeps of every chain, one process per chain, `cores` at a time
            print("bench.py --table: C3 CPU, 96 chains x 10 sweeps on %d processes" % cores, file=sys.stderr,
                  flush=True)
            from concurrent.futures import ThreadPoolExecutor

            def one(chain):
                dims, beta, seed = chain
                (sec, _), = _ref_wait([_ref_bench(3, dims, beta, seed, 10)])
                return sec

            t0 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as pool:
                secs = list(pool.map(one, chains))
            wall = time.perf_counter() - t0
            cpu_sec = sum(secs)
            row["cpu"] = dict(host, kind="reference", build="suncpp -g -O2, FP64, RANLUX3", sweeps_timed=10,
                              one_core_ensemble_sweeps_per_s=10.0 / cpu_sec,
                              one_core_link_updates_per_s=40.0 * V / cpu_sec,
                              one_core_note="sum of the single-threaded processes' own timed seconds, "
                                            "measured while %d ran at once" % cores,
                              all_core_processes=cores, all_core_link_updates_per_s=40.0 * V / wall,
                              all_core_note="wall time of the whole pool, process starts and cold starts "
                                            "included",
                              all_core_wall_s=wall)
        print(json.dumps(row), flush=True)


def copy_calibration(mib=256, repeats=20):
    """A plain copy of known size: each launch reads and writes exactly
    `mib` MiB, so its FETCH_SIZE / WRITE_SIZE per dispatch give the factors
    that turn those counters into HBM bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(repeats):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / repeats
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": repeats + 1,
                      "copy_GBs": 2.0 * n * 4 / (ms * 1e-3) / 1e9}))


def c4(args):
    """The learned-proposal workload (SU(2) 8^3 x 4, beta 2.2) as three child
    processes: heatbath training configurations, forward-KL training (PyTorch
    on the GPU), then the plain and learned chains.  This process never
    touches the GPU; it relays the children's output and prints the chains'
    JSON line last."""
    import tempfile

    work = tempfile.mkdtemp(prefix="qcd_c4_")
    data, params = os.path.join(work, "data.npy"), os.path.join(work, "params.npy")
    py = [sys.executable, "-u", "-m", "qcdgpu_amd.c4"]
    steps = [
        py + ["gen", "--out", data, "--n", str(args.c4_train)],
        py + ["train", "--data", data, "--out", params, "--n", str(args.c4_train), "--steps", str(args.c4_train_steps)],
        py + ["chain", "--params", params, "--steps", str(args.c4_steps)],
    ]
    last = None
    for cmd in steps:
        p = subprocess.Popen(cmd, cwd=REPO, stdout=subprocess.PIPE, text=True)
        for line in p.stdout:
            line = line.rstrip("\n")
            print(line, file=sys.stderr, flush=True)
            last = line
        if p.wait() != 0:
            sys.exit("bench.py --c4: step %s failed" % cmd[4])
    for f in (data, params):
        if os.path.exists(f):
            os.remove(f)
    os.rmdir(work)
    print(last, flush=True)


def c3(args):
    """The batched-ensemble workload: 96 SU(3) chains, (N_s, N_t) drawn from
    {12,16,20,24} x {4,6,8} with seed 54, beta_i = 5.5 + 0.7 i / 95, FP32
    18-real, RANLUX3 per-chain streams, one step = 1 heatbath + 4
    overrelaxation sweeps of every chain with plaquette and Polyakov
    records, captured once as a graph after the autotuner picked each size
    bucket's kernel shape and unit placement."""
    import numpy as np

    import qcdgpu_amd as Q

    chains = _c3_chains()
    V = sum(int(np.prod(c[0])) for c in chains)
    with Q.Batch(chains, precision="fp32", prng="RANLUX", max_measurements=args.steps + args.warmup + 4) as b:
        b.init("hot")
        tuning = b.autotune(repeats=4)
        b.set_merge(not args.c3_per_bucket)
        b.capture_cycle(n_hb=1, n_or=N_OR)
        b.launch_cycle(args.warmup)
        b.sync()
        ms = b.time_cycle(args.steps)
        b.sync()
    ms_step = ms / args.steps
    # 4V link updates per sweep at 9 x 72 B (+ 2 x 112 B of RANLUX state in
    # the heatbath), plaquette and Polyakov 5V x 72 B
    step_bytes = 4.0 * V * ((9 * 72 + 2 * 112) + N_OR * 9 * 72) + 5.0 * V * 72
    gbs = step_bytes / (ms_step * 1e-3) / 1e9
    print(json.dumps({
        "metric": "SU(3) link updates/sec, batched ensemble",
        "value": 5.0 * 4.0 * V / (ms_step * 1e-3),
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "96 SU(3) chains N_s in {12,16,20,24}, N_t in {4,6,8}, beta 5.5-6.2, RANLUX3; "
                               "1 CM heatbath + 4 overrelaxation sweeps + plaquette + Polyakov per step",
                   "sites": V, "tuning": tuning,
                   "launches": "one per size bucket and half-step" if args.c3_per_bucket
                               else "one per half-step and kernel shape for all chains"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_step": step_bytes},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=450)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--storage", type=int, default=18, choices=(18, 12))
    ap.add_argument("--stream", default="reference", choices=("reference", "fast"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--xcd-map", type=int, default=None, choices=(0, 1, 2, 3),
                    help="dispatch order of the update kernels (default: the library's)")
    ap.add_argument("--blocks-per-cu", type=int, default=None,
                    help="cap on resident update workgroups per CU (default: the library's, none)")
    ap.add_argument("--cycles-per-graph", type=int, default=None,
                    help="cycles recorded per graph, measurements overlapped with the next cycle (default: the "
                         "library's, 1 = one cycle per graph)")
    ap.add_argument("--update-shape", type=int, default=None, choices=(1, 3),
                    help="wavefronts per 64-link unit of the update kernels (default: the library's, 3)")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--c3-per-bucket", action="store_true",
                    help="with --c3: one launch per size bucket and half-step instead of one for all chains")
    ap.add_argument(
        "--copy-calibration",
        action="store_true",
        help="only run a plain 256 MiB device-to-device copy 20 times (the known-size "
        "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)",
    )
    ap.add_argument("--c4", action="store_true",
                    help="run the learned-proposal workload instead (SU(2) 8^3x4, beta 2.2; several minutes)")
    ap.add_argument("--c4-train", type=int, default=20000, help="heatbath training configurations")
    ap.add_argument("--c4-train-steps", type=int, default=1000)
    ap.add_argument("--c4-steps", type=int, default=4000, help="steps of each chain")
    ap.add_argument("--c3", action="store_true", help="run the batched-ensemble workload (96 chains) instead")
    ap.add_argument("--table", action="store_true",
                    help="heatbath-only CPU (suncpp) and GPU rows of SU(2) 8^4, SU(3) 32^3x8, 64^4 and the 96-chain "
                         "ensemble, one JSON line each")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.c4:
        c4(args)
        return
    if args.c3:
        c3(args)
        return
    if args.table:
        baseline_table(args)
        return
    if args.gpus != 1:
        print(
            "bench.py: this build measures one GPU; --gpus %d is not supported "
            "(multi-GPU sharding is a design note only)" % args.gpus,
            file=sys.stderr,
        )
        sys.exit(2)

    import qcdgpu_amd as Q

    V = 1
    for d in DIMS:
        V *= d
    lat = Q.Lattice(
        sun=3,
        dims=DIMS,
        beta=BETA,
        precision="fp32",
        storage_reals=args.storage,
        prng="XOR128",
        stream=args.stream,
        randseries=54,
        max_measurements=args.steps + args.warmup + args.profile_steps + 8,
    )
    k_cycles = args.cycles_per_graph or lat.tuning()["cycles_per_graph"]
    lat.tune(xcd_map=args.xcd_map, blocks_per_cu=args.blocks_per_cu, update_shape=args.update_shape,
             cycles_per_graph=1)
    lat.init("cold")
    lat.prng_init()
    # Single-cycle graph first: its step time scales the per-kernel event
    # durations below (no measurement overlapped with the next cycle there).
    lat.capture_cycle(n_hb=1, n_or=N_OR, plaquette=True, polyakov=True)
    lat.launch_cycle(min(args.warmup, 20))
    lat.sync()
    n_single = max(1, min(args.steps, 150))
    ms_single = lat.time_cycle(n_single)
    # The timed workload: K cycles per graph (each cycle's measurement
    # beside the next cycle's first half-steps).
    lat.tune(cycles_per_graph=k_cycles)
    lat.capture_cycle(n_hb=1, n_or=N_OR, plaquette=True, polyakov=True)
    k_captured = lat.tuning()["captured_cycles_per_graph"]
    lat.launch_cycle(args.warmup)
    lat.sync()
    # Timed region: `steps` replays of the captured cycle, HIP events on the
    # lattice stream around the whole region (synchronised on both sides).
    ms = lat.time_cycle(args.steps)
    updates = 5.0 * 4.0 * V * args.steps
    value = updates / (ms * 1e-3)

    prof = lat.profile_cycle(max(1, min(args.profile_steps, args.steps)))
    link_bytes = args.storage * 4
    bytes_per_launch = 9.0 * link_bytes * (V // 2)
    # heatbath half-step: the same links plus a 16-byte stream state read and
    # written per updated link
    hb_bytes = bytes_per_launch + 2.0 * 16 * (V // 2)
    # Kernel durations: HIP events around each kernel of an eagerly launched
    # cycle exclude the gaps between the graph's kernels; scaled so that one
    # cycle's kernels sum to the step of the single-cycle graph, they are the
    # per-kernel share of its timed region (what a kernel trace of those
    # replays reports, which counts each gap into the kernel before it).
    ms_step = ms / args.steps
    n_hb_launch, n_or_launch = 8, 8 * N_OR - 3   # three OR launches of the last sweep carry the plaquette
    event_sum_us = n_hb_launch * prof["hb_us"] + n_or_launch * prof["or_us"] + prof["meas_us"]
    gap_scale = 1000.0 * (ms_single / n_single) / event_sum_us
    or_us, hb_us = prof["or_us"] * gap_scale, prof["hb_us"] * gap_scale
    achieved = bytes_per_launch / (or_us * 1e-6) / 1e9
    hb_achieved = hb_bytes / (hb_us * 1e-6) / 1e9
    # whole step: 4V link updates of each sweep at 9 s (+ 2 x 16 B of stream
    # state for the heatbath), plaquette 4V s and Polyakov V s
    step_bytes = 4.0 * V * ((9 * link_bytes + 32) + N_OR * 9 * link_bytes) + 5.0 * V * link_bytes
    step_achieved = step_bytes / (ms_step * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.storage == 18:
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    meas = lat.measurements()
    lat.close()

    line = {
        "metric": "SU(3) link updates/sec + % HBM roofline",
        "value": value,
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "SU(3) Wilson 32^3x8 beta=5.9: 1 CM heatbath + 4 overrelaxation sweeps + plaquette + Polyakov per step",
            "lattice": list(DIMS),
            "beta": BETA,
            "link_storage": "fp32 %d-real" % args.storage,
            "reunitarisation": "fp64, fused into every link update",
            "prng": "XOR128 per-site streams (%s stream layout)" % args.stream,
            "nhit": 10,
            "start": "cold",
            "updates_per_step": int(5 * 4 * V),
            "kernel_us": {"heatbath_halfstep": hb_us, "overrelax_halfstep": or_us,
                          "measure_per_step": prof["meas_us"] * gap_scale},
            "kernel_us_events": {"heatbath_halfstep": prof["hb_us"], "overrelax_halfstep": prof["or_us"],
                                 "measure_per_step": prof["meas_us"], "gap_scale": gap_scale},
            "cycles_per_graph": k_captured,
            "ms_per_step_single_cycle_graph": ms_single / n_single,
            "sweeps_per_s": args.steps / (ms * 1e-3),
            "heatbath_roofline": {
                "achieved": hb_achieved,
                "unit": "GB/s",
                "frac": hb_achieved / HBM_PEAK_GBS,
                "algorithmic_bytes_per_launch": hb_bytes,
            },
            "step_roofline": {
                "achieved": step_achieved,
                "unit": "GB/s",
                "frac": step_achieved / HBM_PEAK_GBS,
                "algorithmic_bytes_per_step": step_bytes,
            },
            "plaquette_last": float(meas["plq_total"][-1] / 3.0) if "plq_total" in meas else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_update<float,3,%d,overrelax>" % args.storage,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel, scaled by a plain copy "
                              "(profiles/or_halfstep_su3_32c8_fp32_traffic.txt); a profile constant, not measured "
                              "in this run" if traffic else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "duration_us": or_us,
        },
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
