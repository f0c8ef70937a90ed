"""Builds libqcdhip.so in-tree with hipcc for gfx950."""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libqcdhip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Packed FP32 instructions are disabled for all device code.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # a * b + c is fused into one FMA only where the source writes it as one
    # expression, never across statements by the back end: the rounding then
    # follows the source, so every kernel shape, direction instance and
    # dispatch order that runs the same arithmetic gives the same bits.
    "-ffp-contract=on",
]


HOST_SRC = [os.path.join(PKG, "host", n) for n in ("qcdgpu_main.cpp", "host_check.cpp")]
HOST_BIN = os.path.join(PKG, "qcdgpu_hip")


def build_host(verbose=False):
    """The command-line host program (C++ over the C ABI), next to the
    library it loads through an $ORIGIN run path."""
    cxx = os.environ.get("CXX", "g++")
    tmp = HOST_BIN + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-o", tmp] + HOST_SRC + ["-L" + PKG, "-lqcdhip",
           "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("host build failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, HOST_BIN)
    if verbose:
        print("built", HOST_BIN)
    return HOST_BIN


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, n) for n in ("qcd_hip.hip", "qcd_qcg.cpp", "qcd_flow.hip")]
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + srcs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
