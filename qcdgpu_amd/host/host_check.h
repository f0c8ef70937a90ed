// host_check.h -- the end-of-run CPU check of the host program: the final
// configuration is measured again on the host, in FP64, with the host's own
// code, and every measured quantity's last device value is compared with it
// (the reference's LatticeCheckCPU + CPU_GPU_verification_single/double,
// SU(N)/QCDGPU.cpp:254, suncl/su3cpu.cpp:1616-1704,
// data_analysis/data_analysis.cpp:40-104).
#pragma once
#include <string>
#include <utility>
#include <vector>

struct CheckSetup {
    int group = 3;                 // N of SU(N)
    int n[4] = {4, 4, 4, 4};       // N1..N4
    double beta = 5.9;
    bool single = false;           // single-precision links: tolerance 5e-5, else 1e-11
    int pl_level = 3;              // Polyakov loop: > 0 re / im, > 1 also P2 and P4
    bool wilson = false;
    int wilson_r = 7, wilson_t = 2;
    int fmunu = 0;                 // 0 off, 1 H field (xy, xz, yz), 2 E field (xt, yt, zt)
    int fgen[2] = {3, 8};          // generators (Gell-Mann lambda_k, or Pauli sigma_k for SU(2))
    double phi = 0.0, omega = 0.0; // twisted boundary
};

struct CheckResult {
    // (name, device last value, host value) for every compared quantity, in
    // the order of the results file; Fmunu variances (name, host value).
    std::vector<std::pair<std::string, std::pair<double, double>>> rows;
    std::vector<std::pair<std::string, double>> fmunu_var;
    std::vector<std::string> failed;
    bool passed = true;
};

// Host measurement of `links` (canonical layout, FP64:
// links[((site*4 + dir)*N + row)*N + col][re, im], site = y + N2 (z + N3 (t
// + N4 x))) and comparison with the device's last values `gpu_last` (name ->
// value, the names of the results-file columns).  Quantities absent from
// gpu_last are not compared.
CheckResult check_configuration(const CheckSetup& s, const std::vector<double>& links,
                                const std::vector<std::pair<std::string, double>>& gpu_last);

// The results-file block of the check: CPU / GPU last lines and the verdict.
std::string check_report(const CheckSetup& s, const CheckResult& r);
