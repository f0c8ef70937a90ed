"""Pins the oracle's generators (the checker for the GPU PRNGs).

Anchors: glibc rand sequence (fixture), the reference's own CPU twins run from
its sources (RANLUX RL_produce_one_CPU, PM PM_produce_one_CPU; fixtures by
tests/golden/make_golden.py), the reference's self-check design (GPU
instance 0 == CPU twin, random.cpp:339-402), Park & Miller's published
known answer, and Marsaglia's xorshift128 recurrence re-implemented here
independently.
"""
import ctypes
import os

import numpy as np
import pytest


def _f32(golden, name):
    return np.fromfile(os.path.join(golden, name), dtype=np.float32)


@pytest.mark.parametrize("seed", [54, 1, 7])
def test_glibc_rand_matches_fixture(O, golden, seed):
    ref = np.load(os.path.join(golden, "glibc_rand_%d.npy" % seed))
    assert (O.rand_fill(seed, len(ref)).astype(np.int64) == ref).all()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    assert [libc.rand() for _ in range(16)] == list(ref[:16])


@pytest.mark.parametrize("lux", [0, 1, 2, 3, 4])
def test_ranlux_cpu_stream_bit_exact(O, golden, lux):
    ref = _f32(golden, "RANLUX%d_seed54.f32" % lux)
    got = O.ranlux_cpu_stream(54, O.ranlux_nskip(lux), len(ref))
    assert (got.view(np.uint32) == ref.view(np.uint32)).all()


@pytest.mark.parametrize("seed", [1, 7])
def test_ranlux3_other_seeds(O, golden, seed):
    ref = _f32(golden, "RANLUX3_seed%d.f32" % seed)
    got = O.ranlux_cpu_stream(seed, 223, len(ref))
    assert (got.view(np.uint32) == ref.view(np.uint32)).all()


@pytest.mark.parametrize("lux", [0, 1, 2, 3, 4])
def test_ranlux_table_instance0_equals_cpu_twin(O, golden, lux):
    """The float4 kernel's instance 0 must reproduce the scalar CPU stream
    (PRNG::check, random.cpp:339-402), across many 24-blocks and skips."""
    ref = _f32(golden, "RANLUX%d_seed54.f32" % lux)
    table = O.ranlux_seed_table(54, 5)
    q = O.ranlux_produce(table, len(ref) // 4, O.ranlux_nskip(lux))
    assert (q[:, 0, :].ravel().view(np.uint32) == ref.view(np.uint32)).all()


def test_ranlux_produce_resumes_across_calls(O):
    """Two produce() calls equal one call of the combined length (the table
    carries the ring position between kernel launches)."""
    t1 = O.ranlux_seed_table(3, 4)
    t2 = t1.copy()
    a = O.ranlux_produce(t1, 10, 223)
    b = O.ranlux_produce(t1, 7, 223)
    c = O.ranlux_produce(t2, 17, 223)
    assert (np.concatenate([a, b]).view(np.uint32) == c.view(np.uint32)).all()


def test_ranlux_range_and_distribution(O):
    x = O.ranlux_cpu_stream(54, 223, 200000)
    assert x.min() > 0.0 and x.max() < 1.0
    assert abs(x.mean() - 0.5) < 0.003
    assert abs(x.var() - 1.0 / 12) < 0.002


def test_pm_cpu_twin_fixture(O, golden):
    for seed in (54, 1):
        ref = _f32(golden, "PM_seed%d.f32" % seed)
        O.lib().orc_srand(seed)
        s0 = O.lib().orc_rand() % 2147483647
        got = O.pm_cpu_stream(s0, len(ref))
        assert (got.view(np.uint32) == ref.view(np.uint32)).all()


def test_pm_park_miller_known_answer(O):
    """Park & Miller (1988): from seed 1 the 10000th value is 1043618065
    (prngcl_pm.cl:71 comment)."""
    s = 1
    for _ in range(10000):
        s = O.lib().orc_pm_step(s)
    assert s == 1043618065


def test_pm_kernel_lane_x_equals_cpu_twin(O):
    """Lane x of instance 0 is the CPU stream; the kernel's float division by
    (float)(2^31-1) = 2^31 differs from the twin's double division only in the
    last float bit."""
    t = O.pm_seed_table(54, 3)
    seed0 = int(t[0, 0])
    q = O.pm_produce(t.copy(), 500)
    cpu = O.pm_cpu_stream(seed0, 500)
    d = np.abs(q[:, 0, 0].astype(np.float64) - cpu.astype(np.float64))
    assert d.max() <= 2.0 ** -23


def _xorshift128_textbook(x, y, z, w, n):
    """Marsaglia, J. Stat. Soft. 8(14) 2003, p.5: t=x^(x<<11); x=y; y=z; z=w;
    return w = w^(w>>19)^(t^(t>>8))."""
    out = []
    M = 0xFFFFFFFF
    for _ in range(n):
        t = (x ^ (x << 11)) & M
        x, y, z = y, z, w
        w = (w ^ (w >> 19) ^ (t ^ (t >> 8))) & M
        out.append(w)
    return out


def test_xor128_marsaglia_published_seed(O):
    """Marsaglia's default state (123456789, 362436069, 521288629, 88675123)
    gives 3701687786 as the first output."""
    s = (ctypes.c_uint32 * 4)(123456789, 362436069, 521288629, 88675123)
    first = O.lib().orc_xor128_step(s)
    assert first == 3701687786
    ref = _xorshift128_textbook(123456789, 362436069, 521288629, 88675123, 1000)
    s = (ctypes.c_uint32 * 4)(123456789, 362436069, 521288629, 88675123)
    got = [O.lib().orc_xor128_step(s) for _ in range(1000)]
    assert got == ref


def test_xor128_produce_layout_and_scaling(O):
    n = 5
    t = O.xor128_seed_table(54, n)
    libc_vals = O.rand_fill(54, 4 * n)
    assert (t.ravel().astype(np.int64) == libc_vals.astype(np.int64)).all()
    q = O.xor128_produce(t.copy(), 3)
    for g in range(n):
        w = _xorshift128_textbook(*[int(v) for v in O.xor128_seed_table(54, n)[g]], 12)
        exp = np.array(w, dtype=np.float64).astype(np.float32) / np.float32(4294967296.0)
        assert (q[:, g, :].ravel().view(np.uint32) == exp.astype(np.float32).view(np.uint32)).all()
