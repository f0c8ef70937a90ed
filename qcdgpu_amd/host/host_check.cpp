// host_check.cpp -- host FP64 measurement of the final configuration and
// its comparison with the device's last measurement (the reference's
// LatticeCheckCPU, SU(N)/QCDGPU.cpp:254 and suncl/suncpu.cpp:41-56, with the
// measurements of suncl/su3cpu.cpp:1155-1522 / su2cpu.cpp and the rule of
// CPU_GPU_verification_single/double, data_analysis/data_analysis.cpp:76-104:
// a quantity fails when both values exceed 1e-6 (single precision) or 1e-12
// (double) in magnitude and |1 - cpu/gpu| exceeds 5e-5 or 1e-11).
//
// Normalisations are those of the device records as the host program
// prints them: plaquette sums over 3V, Polyakov sums over V_s N (P2 over
// V_s N^2, P4 over V_s N^4), F_munu sums over V, Wilson loop sums over 3V.
// (The reference's CPU Wilson loop divides by N V, su3cpu.cpp:1521, which
// equals its device normalisation 3V only for SU(3); here both sides use 3V.)
#include "host_check.h"

#include <cmath>
#include <complex>
#include <cstdio>

namespace {

using cd = std::complex<double>;

struct Mat {
    cd a[3][3];
};

Mat mul(const Mat& x, const Mat& y, int n)
{
    Mat r{};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cd s = 0.0;
            for (int k = 0; k < n; ++k) s += x.a[i][k] * y.a[k][j];
            r.a[i][j] = s;
        }
    return r;
}

Mat dag(const Mat& x, int n)
{
    Mat r{};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) r.a[i][j] = std::conj(x.a[j][i]);
    return r;
}

cd trace(const Mat& x, int n)
{
    cd s = 0.0;
    for (int i = 0; i < n; ++i) s += x.a[i][i];
    return s;
}

Mat unit(int n)
{
    Mat r{};
    for (int i = 0; i < n; ++i) r.a[i][i] = 1.0;
    return r;
}

// Gell-Mann lambda_k (k = 1..8) or Pauli sigma_k (k = 1..3).
Mat generator(int n, int k)
{
    const cd I(0.0, 1.0);
    Mat g{};
    if (n == 2) {
        if (k == 1) g.a[0][1] = g.a[1][0] = 1.0;
        else if (k == 2) g.a[0][1] = -I, g.a[1][0] = I;
        else g.a[0][0] = 1.0, g.a[1][1] = -1.0;
        return g;
    }
    switch (k) {
        case 1: g.a[0][1] = g.a[1][0] = 1.0; break;
        case 2: g.a[0][1] = -I; g.a[1][0] = I; break;
        case 3: g.a[0][0] = 1.0; g.a[1][1] = -1.0; break;
        case 4: g.a[0][2] = g.a[2][0] = 1.0; break;
        case 5: g.a[0][2] = -I; g.a[2][0] = I; break;
        case 6: g.a[1][2] = g.a[2][1] = 1.0; break;
        case 7: g.a[1][2] = -I; g.a[2][1] = I; break;
        default: {
            const double r = 1.0 / std::sqrt(3.0);
            g.a[0][0] = r;
            g.a[1][1] = r;
            g.a[2][2] = -2.0 * r;
        }
    }
    return g;
}

struct Field {
    const CheckSetup& s;
    const std::vector<double>& links;
    int N;
    cd tw[3];
    bool twisted;

    Field(const CheckSetup& st, const std::vector<double>& l) : s(st), links(l), N(st.group)
    {
        // the twisted boundary: Y links at x = N1 - 1 are read as diag(tw) U
        // (su3_matrix_memory.cl:57-98, su2_matrix_memory.cl:56-88)
        const double ang[3] = {N == 3 ? 0.5 * (s.phi + s.omega) : 0.5 * s.phi,
                               N == 3 ? 0.5 * (s.omega - s.phi) : -0.5 * s.phi, N == 3 ? -s.omega : 0.0};
        for (int r = 0; r < 3; ++r) tw[r] = std::polar(1.0, ang[r]);
        twisted = (N == 3) ? (s.phi != 0.0 || s.omega != 0.0) : s.phi != 0.0;
    }
    static int wrap(int v, int n) { return ((v % n) + n) % n; }
    Mat link(int x, int y, int z, int t, int dir) const
    {
        x = wrap(x, s.n[0]);
        y = wrap(y, s.n[1]);
        z = wrap(z, s.n[2]);
        t = wrap(t, s.n[3]);
        const size_t site = (size_t)y + (size_t)s.n[1] * ((size_t)z + (size_t)s.n[2] * ((size_t)t + (size_t)s.n[3] * x));
        const double* p = &links[((site * 4 + dir) * N * N) * 2];
        Mat m{};
        for (int r = 0; r < N; ++r)
            for (int c = 0; c < N; ++c) m.a[r][c] = cd(p[2 * (r * N + c)], p[2 * (r * N + c) + 1]);
        if (twisted && dir == 1 && x == s.n[0] - 1)
            for (int r = 0; r < N; ++r)
                for (int c = 0; c < N; ++c) m.a[r][c] *= tw[r];
        return m;
    }
    Mat link(const int c[4], int dir) const { return link(c[0], c[1], c[2], c[3], dir); }
};

// U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+
Mat plaquette(const Field& f, const int c[4], int a, int b)
{
    int ca[4] = {c[0], c[1], c[2], c[3]}, cb[4] = {c[0], c[1], c[2], c[3]};
    ca[a] += 1;
    cb[b] += 1;
    const int n = f.N;
    return mul(mul(f.link(c, a), f.link(ca, b), n), dag(mul(f.link(c, b), f.link(cb, a), n), n), n);
}

// Product of `len` links of direction dir starting at c.
Mat line(const Field& f, int c0[4], int dir, int len)
{
    int c[4] = {c0[0], c0[1], c0[2], c0[3]};
    Mat p = unit(f.N);
    for (int k = 0; k < len; ++k) {
        p = mul(p, f.link(c, dir), f.N);
        c[dir] += 1;
    }
    return p;
}

}  // namespace

CheckResult check_configuration(const CheckSetup& s, const std::vector<double>& links,
                                const std::vector<std::pair<std::string, double>>& gpu_last)
{
    const Field f(s, links);
    const int N = s.group;
    const double V = (double)s.n[0] * s.n[1] * s.n[2] * s.n[3];
    const double Vs = V / s.n[3];
    double spat = 0.0, temp = 0.0, pl[4] = {0, 0, 0, 0}, wil = 0.0;
    double fsum[2][3][2] = {}, fsq[2][3][2] = {};
    const int planes[2][3][2] = {{{0, 1}, {0, 2}, {1, 2}}, {{0, 3}, {1, 3}, {2, 3}}};
    const Mat G[2] = {generator(N, s.fgen[0]), generator(N, s.fgen[1])};
    int c[4];
    for (c[0] = 0; c[0] < s.n[0]; ++c[0])
        for (c[1] = 0; c[1] < s.n[1]; ++c[1])
            for (c[2] = 0; c[2] < s.n[2]; ++c[2])
                for (c[3] = 0; c[3] < s.n[3]; ++c[3]) {
                    for (int a = 0; a < 3; ++a)
                        for (int b = a + 1; b < 4; ++b) {
                            const Mat P = plaquette(f, c, a, b);
                            const double r = trace(P, N).real();
                            (b == 3 ? temp : spat) += r;
                            if (s.fmunu) {
                                const int fld = s.fmunu - 1;
                                for (int pl_i = 0; pl_i < 3; ++pl_i) {
                                    if (planes[fld][pl_i][0] != a || planes[fld][pl_i][1] != b) continue;
                                    for (int g = 0; g < 2; ++g) {
                                        const cd t = trace(mul(P, G[g], N), N);
                                        // F = -i Tr(P G) / 2 (SU(2): the imaginary part is 0)
                                        const double re = 0.5 * t.imag(), im = (N == 2) ? 0.0 : -0.5 * t.real();
                                        fsum[g][pl_i][0] += re;
                                        fsum[g][pl_i][1] += im;
                                        fsq[g][pl_i][0] += re * re;
                                        fsq[g][pl_i][1] += im * im;
                                    }
                                }
                            }
                        }
                    if (s.wilson) {
                        for (int d = 0; d < 3; ++d) {
                            int ct[4] = {c[0], c[1], c[2], c[3]}, cr[4] = {c[0], c[1], c[2], c[3]};
                            ct[3] += s.wilson_t;
                            cr[d] += s.wilson_r;
                            const Mat B = line(f, c, 3, s.wilson_t), R = line(f, ct, d, s.wilson_r);
                            const Mat T = line(f, cr, 3, s.wilson_t), L = line(f, c, d, s.wilson_r);
                            wil += trace(mul(mul(B, R, N), dag(mul(L, T, N), N), N), N).real();
                        }
                    }
                }
    if (s.pl_level > 0)
        for (c[0] = 0; c[0] < s.n[0]; ++c[0])
            for (c[1] = 0; c[1] < s.n[1]; ++c[1])
                for (c[2] = 0; c[2] < s.n[2]; ++c[2]) {
                    c[3] = 0;
                    const cd t = trace(line(f, c, 3, s.n[3]), N);
                    const double a2 = std::norm(t);
                    pl[0] += t.real();
                    pl[1] += t.imag();
                    pl[2] += a2;
                    pl[3] += a2 * a2;
                }

    const double bet = s.beta / N;
    const double ps = spat / (3 * V), pt = temp / (3 * V);
    std::vector<std::pair<std::string, double>> cpu = {
        {"Plq_spat", ps}, {"Plq_temp", pt}, {"Plq_total", 0.5 * (ps + pt)},
        {"S_spat", bet * (N - ps)}, {"S_temp", bet * (N - pt)}, {"S_total", 0.5 * (bet * (N - ps) + bet * (N - pt))},
    };
    if (s.wilson) cpu.push_back({"Wilson_loop", wil / (3 * V)});
    if (s.pl_level > 0) {
        cpu.push_back({"Polyakov_loop", pl[0] / (Vs * N)});
        cpu.push_back({"Polyakov_loop_im", pl[1] / (Vs * N)});
    }
    if (s.pl_level > 1) {
        cpu.push_back({"Polyakov_loop_P2", pl[2] / (Vs * N * N)});
        cpu.push_back({"Polyakov_loop_P4", pl[3] / (Vs * N * N * N * N)});
    }
    CheckResult res;
    if (s.fmunu) {
        const char* ph[3] = {"xy", "xz", "yz"};
        const char* pe[3] = {"xt", "yt", "zt"};
        const char* ri[2] = {"re", "im"};
        double absq[2][2] = {};
        for (int g = 0; g < 2; ++g)
            for (int p = 0; p < 3; ++p)
                for (int k = 0; k < 2; ++k) {
                    const std::string name = std::string("Fmunu_") + (s.fmunu == 1 ? ph[p] : pe[p]) + "_" +
                                             std::to_string(s.fgen[g]) + "_" + ri[k];
                    const double m = fsum[g][p][k] / V;
                    cpu.push_back({name, m});
                    res.fmunu_var.push_back({name, fsq[g][p][k] / V - m * m});
                    absq[g][k] += m * m;
                }
        for (int g = 0; g < 2; ++g)
            for (int k = 0; k < 2; ++k)
                cpu.push_back({std::string("Fmunu_abs_") + std::to_string(s.fgen[g]) + "_" + ri[k], std::sqrt(absq[g][k])});
    }
    const double thr = s.single ? 1e-6 : 1e-12, tol = s.single ? 5e-5 : 1e-11;
    for (const auto& q : cpu) {
        const std::pair<std::string, double>* g = nullptr;
        for (const auto& e : gpu_last)
            if (e.first == q.first) g = &e;
        if (!g) continue;
        const double a = g->second, b = q.second;
        res.rows.push_back({q.first, {a, b}});
        if (std::fabs(a) > thr && std::fabs(b) > thr && !(std::fabs(1.0 - b / a) <= tol)) {
            res.failed.push_back(q.first);
            res.passed = false;
        }
    }
    return res;
}

std::string check_report(const CheckSetup& s, const CheckResult& r)
{
    std::string out;
    char line[256];
    const char* sep = " ***************************************************\n";
    out += sep;
    for (const auto& q : r.rows) {
        snprintf(line, sizeof(line), " GPU last %-20s: % 16.13e\n", q.first.c_str(), q.second.first);
        out += line;
        snprintf(line, sizeof(line), " CPU last %-20s: % 16.13e\n", q.first.c_str(), q.second.second);
        out += line;
    }
    if (!r.fmunu_var.empty()) {
        out += sep;
        for (const auto& q : r.fmunu_var) {
            snprintf(line, sizeof(line), " CPU l.varnc %-17s: % 16.13e\n", q.first.c_str(), q.second);
            out += line;
        }
    }
    out += sep;
    snprintf(line, sizeof(line), " tolerance %g relative (%s precision links)\n", s.single ? 5e-5 : 1e-11,
             s.single ? "single" : "double");
    out += line;
    out += r.passed ? " *** Verification successfully passed! *************\n"
                    : " --- Verification failed! --------------------------\n";
    return out;
}
