"""Writes the golden fixtures in tests/golden/ from the reference's own CPU path.

Run in the build container only (it needs the reference tree):

    oracle/build_ref.sh && python tests/golden/make_golden.py

Sources of each fixture:
  glibc_rand_<seed>.npy  -- glibc srand/rand (the reference seeds every
                            generator through them, random.cpp:235,465,508)
  <GEN>_seed<s>.f32      -- PRNG::produce_CPU of the reference
                            (random.cpp:301-316) via oracle/_ref/ref_harness
  sweep_<tag>_s<k>.npy   -- links after k sweeps of the reference suncpp
                            (sun_update.h:80-117, order of suncpp.h:159-168),
                            float64 [site][dir][row][col][re,im]
  sweep_<tag>_plq.txt    -- plqConf (Plq.h:47-94): sweep, spat, temp, total
"""
import ctypes
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

# (tag, sun, dims, beta, nhit, ints, generator, randseries, sweeps kept)
SWEEPS = [
    ("su3_gid_4444", 3, (4, 4, 4, 4), 5.9, 10, 2, "RANLUX3", 54, (1, 2)),
    ("su3_gid_6424", 3, (6, 4, 2, 4), 5.9, 10, 2, "RANLUX3", 54, (1,)),
    ("su3_cold_rl3_6424", 3, (6, 4, 2, 4), 5.9, 10, 1, "RANLUX3", 54, (1, 2)),
    ("su3_cold_pm_4444", 3, (4, 4, 4, 4), 6.0, 10, 1, "PM", 7, (1,)),
    ("su2_gid_4444", 2, (4, 4, 4, 4), 2.3, 10, 2, "RANLUX3", 54, (1, 2)),
    ("su2_cold_rl3_6424", 2, (6, 4, 2, 4), 2.3, 10, 1, "RANLUX3", 54, (1, 2)),
    ("su2_cold_rl0_4444", 2, (4, 4, 4, 4), 1.7, 2, 1, "RANLUX0", 1, (1,)),
]

PRNG_STREAMS = [("RANLUX%d" % l, s, 2048) for l in range(5) for s in (54,)] + [
    ("RANLUX3", 1, 2048),
    ("RANLUX3", 7, 2048),
    ("PM", 54, 1024),
    ("PM", 1, 1024),
]


def run(args):
    subprocess.run([HARNESS] + [str(a) for a in args], check=True)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build oracle/_ref/ref_harness first (oracle/build_ref.sh)")
    libc = ctypes.CDLL("libc.so.6")
    for seed in (54, 1, 7):
        libc.srand(seed)
        vals = np.array([libc.rand() for _ in range(256)], dtype=np.int64)
        np.save(os.path.join(HERE, "glibc_rand_%d.npy" % seed), vals)

    for gen, seed, count in PRNG_STREAMS:
        run(["prng", gen, seed, count, os.path.join(HERE, "%s_seed%d.f32" % (gen, seed))])

    tmp = tempfile.mkdtemp()
    try:
        for tag, sun, dims, beta, nhit, ints, gen, seed, keep in SWEEPS:
            prefix = os.path.join(tmp, tag)
            run(["sweep", sun, *dims, repr(beta), nhit, ints, gen, seed, max(keep), prefix])
            V = int(np.prod(dims))
            for k in keep:
                a = np.fromfile("%s_s%d.bin" % (prefix, k), dtype=np.float64)
                a = a.reshape(V, 4, sun, sun, 2)
                np.save(os.path.join(HERE, "sweep_%s_s%d.npy" % (tag, k)), a)
            shutil.copy(prefix + "_plq.txt", os.path.join(HERE, "sweep_%s_plq.txt" % tag))
    finally:
        shutil.rmtree(tmp)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
