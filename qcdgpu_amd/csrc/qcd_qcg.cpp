// qcd_qcg.cpp -- the reference's .qcg state file, written and read over the
// C ABI of libqcdhip.so (host code only).
//
// Layout: model::lattice_save_state, lattice_make_bin_header,
// lattice_load_bin_header and lattice_load_state (SU(N)/suncl/suncl.cpp:
// 1857-2079); section sizes from model::model_lattice_init
// (suncl.cpp:2266-2298) through GPU::buffer_size_align
// (clinterface/clinterface.cpp:1580-1590) at the 64-item work-group cap the
// reference driver sets (QCDGPU.cpp:74); table planes from lattice_store_3 /
// lattice_store_2 (suncl/su3_matrix_memory.cl:150-176,
// su2_matrix_memory.cl:128-145) and lattice_reconstruct3 / 2
// (suncl/su3cl.cl:137-163, su2cl.cl:63-79).  After the table the writer
// appends an extension the reference never reads: "QCDGPUX1", a uint64 byte
// count and the exact state image of qcd_lattice_state_save.

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/qcd_hip.h"
#include "qcd_internal.h"

namespace {

constexpr int kHeadWords = 64;      // BIN_HEADER_SIZE (suncl.cpp:60)
constexpr uint64_t kAlign = 64;     // buffer_size_align step
constexpr char kExtMagic[8] = {'Q', 'C', 'D', 'G', 'P', 'U', 'X', '1'};
const char* const kPrefix = "QCDGPU";   // bin_prefix (suncl.cpp:1946)
const char* const kVersion = "1.5";     // version (QCDGPU.cpp:42)

int err(int code, const std::string& msg) { return qcdhip::set_error(code, msg.c_str()); }

uint64_t align64(uint64_t x) { return x ? (1 + (x - 1) / kAlign) * kAlign : 0; }

// model::convert_str_uint (suncl.cpp:709-716): bytes off..off+3, little end first.
uint32_t str_word(const char* s, size_t off)
{
    const size_t n = strlen(s);
    uint32_t r = 0;
    for (unsigned i = 0; i < 4; ++i)
        if (off + i < n) r += (uint32_t)(unsigned char)s[off + i] << (8 * i);
    return r;
}

// PRNG::convert_generator_to_uint (random.cpp:110-122) and RL_nskip
// (random.cpp:66,192-196).
uint32_t generator_code(int prng, int lux)
{
    if (prng == QCD_PRNG_XOR128) return 1;
    if (prng == QCD_PRNG_PM) return 9;
    return 2u + (uint32_t)lux;   // RANLUX0..RANLUX4
}
uint32_t rl_nskip(int prng, int lux)
{
    static const uint32_t skip[5] = {24, 48, 97, 223, 389};
    return (prng == QCD_PRNG_RANLUX && lux >= 0 && lux <= 4) ? skip[lux] : 24u;
}

struct Layout {
    uint64_t V = 0;
    uint64_t rowsize = 0;       // lattice_table_row_size
    uint64_t table_size = 0;    // 4-vectors
    uint64_t meas_size_F = 0;     // double2 (per-workgroup partials; x7 with F_munu)
    uint64_t energies_size = 0;   // double2 / double per record slot
    uint64_t energies_size_F = 0; // double2: plaquette section (x7 with F_munu)
    uint64_t poly_size = 0;       // double2
    size_t elem = 0;              // bytes per 4-vector: 16 (float4) or 32 (double4)
    int N = 3;
    uint32_t get_plq = 1, get_wilson = 0, pl_level = 0;
    uint32_t f = 0;               // F_munu sections (get_Fmunu or get_F0mu)

    uint64_t body_bytes() const
    {
        uint64_t b = 16 * meas_size_F + 16 * energies_size;
        if (get_plq || f) b += 16 * energies_size_F;
        if (get_wilson) b += 8 * energies_size;
        if (pl_level > 0) b += 16 * poly_size;
        return b + (uint64_t)elem * table_size;
    }
};

Layout make_layout(int sun, const int n[4], int precision, unsigned iter, unsigned get_plq, unsigned get_wilson,
                   unsigned pl_level, unsigned f)
{
    Layout l;
    l.N = sun;
    l.V = (uint64_t)n[0] * n[1] * n[2] * n[3];
    l.rowsize = align64(l.V);
    const uint64_t elements = (sun == 3) ? 12 : 4;   // lattice_group_elements (suncl.cpp:751-753)
    l.table_size = align64(l.rowsize * 4 * elements / 4);
    const uint64_t kF = f ? 7 : 1;   // MODEL_energies_size (suncl.h:221) with F_munu
    l.meas_size_F = kF * align64((align64(l.V) + 63) / 64);   // ceil(exact_row_size / 64)
    const uint64_t it = iter ? iter : 1;
    l.energies_size = align64(it);
    l.energies_size_F = kF * l.energies_size;
    l.f = f ? 1 : 0;
    l.poly_size = align64(it);
    l.elem = (precision == QCD_FP32) ? 16 : 32;
    l.get_plq = get_plq ? 1 : 0;
    l.get_wilson = get_wilson ? 1 : 0;
    l.pl_level = pl_level;
    return l;
}

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

bool put(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

// Reference planes of one link from the canonical row-major complex matrix m.
void to_planes(int N, const double* m, double out[12])
{
    auto re = [&](int r, int c) { return m[2 * (r * N + c)]; };
    auto im = [&](int r, int c) { return m[2 * (r * N + c) + 1]; };
    if (N == 3) {
        const double p[12] = {re(0, 0), re(0, 1), re(0, 2), re(1, 2), im(0, 0), im(0, 1),
                              im(0, 2), im(1, 2), re(1, 0), re(1, 1), im(1, 0), im(1, 1)};
        memcpy(out, p, sizeof(p));
    } else {
        const double p[4] = {re(0, 0), re(0, 1), im(0, 0), im(0, 1)};
        memcpy(out, p, sizeof(p));
    }
}

// Canonical matrix from reference planes (rows 1-2 stored, row 3 =
// conj(row1 x row2) for SU(3); SU(2) row 2 = (-conj u12, conj u11)).
void from_planes(int N, const double* p, double* m)
{
    if (N == 3) {
        const double u[3][2] = {{p[0], p[4]}, {p[1], p[5]}, {p[2], p[6]}};
        const double v[3][2] = {{p[8], p[10]}, {p[9], p[11]}, {p[3], p[7]}};
        for (int c = 0; c < 3; ++c) {
            m[2 * c] = u[c][0];
            m[2 * c + 1] = u[c][1];
            m[2 * (3 + c)] = v[c][0];
            m[2 * (3 + c) + 1] = v[c][1];
        }
        for (int c = 0; c < 3; ++c) {
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            const double re = u[a][0] * v[b][0] - u[a][1] * v[b][1] - (u[b][0] * v[a][0] - u[b][1] * v[a][1]);
            const double im = u[a][0] * v[b][1] + u[a][1] * v[b][0] - (u[b][0] * v[a][1] + u[b][1] * v[a][0]);
            m[2 * (6 + c)] = re;
            m[2 * (6 + c) + 1] = -im;
        }
    } else {
        m[0] = p[0];
        m[1] = p[2];
        m[2] = p[1];
        m[3] = p[3];
        m[4] = -p[1];
        m[5] = p[3];
        m[6] = p[0];
        m[7] = -p[2];
    }
}

struct Header {
    uint32_t w[kHeadWords];
};

// The lattice parameters a header describes (storage, stream layout and
// record capacity are not in it: 18 / 4 reals, reference streams, ITER).
int header_to_params(const Header& h, qcd_params* p, qcd_run_info* info)
{
    const uint32_t* w = h.w;
    if (w[0] != str_word(kPrefix, 0) || w[1] != str_word(kPrefix, 4))
        return err(QCD_E_ARG, "not a QCDGPU state file (prefix)");
    if (w[2] != str_word(kVersion, 0) || w[3] != 0)
        return err(QCD_E_ARG, "state file of another QCDGPU version");
    qcd_params q = {};
    q.sun = (int)w[36];
    if (q.sun != 2 && q.sun != 3) return err(QCD_E_UNSUPPORTED, "state file group SU(" + std::to_string(w[36]) + ")");
    if (w[37] != 4) return err(QCD_E_UNSUPPORTED, "state file with ND = " + std::to_string(w[37]));
    for (int k = 0; k < 4; ++k) {
        if (w[38 + k] != w[42 + k])
            return err(QCD_E_UNSUPPORTED, "state file of a multi-device run (domain differs from the lattice)");
        q.n[k] = (int)w[38 + k];
    }
    if (w[18] == 1) q.precision = QCD_FP32;
    else if (w[18] == 2) q.precision = QCD_FP64;
    else return err(QCD_E_UNSUPPORTED, "state file precision code " + std::to_string(w[18]));
    const uint32_t gen = w[7];
    if (gen == 1) q.prng = QCD_PRNG_XOR128;
    else if (gen == 9) q.prng = QCD_PRNG_PM;
    else if (gen >= 2 && gen <= 6) {
        q.prng = QCD_PRNG_RANLUX;
        q.ranlux_luxury = (int)gen - 2;
    } else {
        return err(QCD_E_UNSUPPORTED, "state file generator code " + std::to_string(gen));
    }
    uint64_t bb = (uint64_t)w[19] | ((uint64_t)w[20] << 32);
    memcpy(&q.beta, &bb, 8);
    double phi, omega;
    const uint64_t pb = (uint64_t)w[21] | ((uint64_t)w[22] << 32), ob = (uint64_t)w[23] | ((uint64_t)w[24] << 32);
    memcpy(&phi, &pb, 8);
    memcpy(&omega, &ob, 8);
    if (!std::isfinite(phi) || !std::isfinite(omega)) return err(QCD_E_ARG, "state file PHI / OMEGA not finite");
    if (w[34] != w[33] && (uint64_t)w[34] != 7ull * w[33])
        return err(QCD_E_ARG, "state file section sizes: energies_size_F is neither 1 nor 7 x energies_size");
    const bool fsec = w[34] != w[33];
    if (w[28] && !fsec) return err(QCD_E_ARG, "state file with get_Fmunu but no F_munu sections");
    q.storage_reals = (q.sun == 3) ? 18 : 4;
    q.nhit = (int)w[15];
    q.nhitpar = 1;
    q.stream_mode = QCD_STREAM_REFERENCE;
    q.randseries = w[6];
    q.max_measurements = (int)(w[13] ? w[13] : 1);
    if (p) *p = q;
    if (info) {
        info->init = w[4];
        info->start = w[5];
        info->prng_counter = w[9];
        info->nav = w[10];
        info->nav_counter = w[11];
        info->niter = w[12];
        info->iter = w[13];
        info->iter_counter = w[14];
        info->wilson_r = w[16];
        info->wilson_t = w[17];
        info->pl_level = w[25];
        info->get_actions = w[29];
        info->get_plaquettes = w[30];
        info->get_wilson = w[31];
        info->fmunu_index1 = w[26];
        info->fmunu_index2 = w[27];
        info->get_fmunu = w[28] ? 1u : 0u;
        info->get_f0mu = (fsec && !w[28]) ? 1u : 0u;
        info->phi = phi;
        info->omega = omega;
    }
    return QCD_OK;
}

// Section sizes come from the header words, as lattice_load_bin_header
// reads them (suncl.cpp:2031-2034).
Layout header_layout(const Header& h, const qcd_params& q)
{
    Layout l = make_layout(q.sun, q.n, q.precision, h.w[13], h.w[30], h.w[31], h.w[25], 0);
    l.meas_size_F = h.w[32];
    l.energies_size = h.w[33];
    l.energies_size_F = h.w[34];
    l.poly_size = h.w[35];
    l.f = (h.w[34] != h.w[33]) ? 1 : 0;
    return l;
}

// Settings the header carries: the F_munu generators of a file with F
// sections (words 26-27) and the twist (words 21-24).
int apply_settings(qcd_lattice* lat, const qcd_run_info& ri)
{
    int rc = qcd_lattice_set_twist(lat, ri.phi, ri.omega);
    if (rc || (!ri.get_fmunu && !ri.get_f0mu)) return rc;
    return qcd_lattice_set_fmunu(lat, (int)ri.fmunu_index1, (int)ri.fmunu_index2);
}

// Parameter fields of an exact image head (see qcd_lattice_state_save):
// magic[8], head bytes, 14 uint32 fields, beta bits; version 2 holds
// prng_instances at byte 128 (version 1 had none: 0).
bool image_params(const unsigned char* head, size_t n, qcd_params* q)
{
    const bool v2 = n >= 8 && memcmp(head, "QCDSTAT2", 8) == 0;
    if (n < 8 + 4 + 14 * 4 + 8 || (!v2 && memcmp(head, "QCDSTAT1", 8) != 0)) return false;
    if (v2 && n < 132) return false;
    uint32_t f[14];
    memcpy(f, head + 12, sizeof(f));
    q->sun = (int)f[0];
    for (int k = 0; k < 4; ++k) q->n[k] = (int)f[1 + k];
    q->precision = (int)f[5];
    q->storage_reals = (int)f[6];
    q->nhit = (int)f[7];
    q->nhitpar = (int)f[8];
    q->prng = (int)f[9];
    q->ranlux_luxury = (int)f[10];
    q->stream_mode = (int)f[11];
    q->randseries = f[12];
    q->max_measurements = (int)f[13];
    memcpy(&q->beta, head + 12 + 56, 8);
    q->prng_instances = 0;
    if (v2) memcpy(&q->prng_instances, head + 128, 4);
    return true;
}

// Start flags of an exact image (after the parameter fields and the
// production counter): bit 0 streams seeded, bit 1 produce() sized as in the
// reference's hot runs.
bool image_start_flags(const unsigned char* head, size_t n, uint32_t* flags)
{
    constexpr size_t at = 8 + 4 + 14 * 4 + 8 + 4;
    if (n < at + 4 || (memcmp(head, "QCDSTAT1", 8) != 0 && memcmp(head, "QCDSTAT2", 8) != 0)) return false;
    memcpy(flags, head + at, 4);
    return true;
}

bool same_params(const qcd_params& a, const qcd_params& b)
{
    bool s = a.sun == b.sun && a.precision == b.precision && a.storage_reals == b.storage_reals &&
             a.nhit == b.nhit && a.nhitpar == b.nhitpar && a.prng == b.prng && a.stream_mode == b.stream_mode &&
             a.randseries == b.randseries && a.max_measurements == b.max_measurements &&
             a.prng_instances == b.prng_instances && memcmp(&a.beta, &b.beta, 8) == 0;
    if (a.prng == QCD_PRNG_RANLUX) s = s && a.ranlux_luxury == b.ranlux_luxury;
    for (int k = 0; k < 4; ++k) s = s && a.n[k] == b.n[k];
    return s;
}

// Opens `path` and reads the header, the layout and the extension position.
int open_state(const char* path, File& fh, Header& h, qcd_params& q, qcd_run_info* info, Layout& lay,
               uint64_t& ext_bytes)
{
    if (!path) return err(QCD_E_ARG, "null path");
    fh.f = fopen(path, "rb");
    if (!fh.f) return err(QCD_E_ARG, std::string("cannot open ") + path + ": " + strerror(errno));
    if (!get(fh.f, h.w, sizeof(h.w))) return err(QCD_E_ARG, std::string(path) + ": short header");
    int rc = header_to_params(h, &q, info);
    if (rc) return rc;
    lay = header_layout(h, q);
    ext_bytes = 0;
    const uint64_t end = sizeof(h.w) + lay.body_bytes();
    if (fseeko(fh.f, 0, SEEK_END) != 0) return err(QCD_E_ARG, "cannot seek in state file");
    const uint64_t size = (uint64_t)ftello(fh.f);
    if (size < end) return err(QCD_E_ARG, std::string(path) + ": truncated (" + std::to_string(size) + " of " +
                                              std::to_string(end) + " bytes)");
    if (size >= end + 16) {
        char magic[8];
        uint64_t n = 0;
        if (fseeko(fh.f, (off_t)end, SEEK_SET) != 0 || !get(fh.f, magic, 8) || !get(fh.f, &n, 8))
            return err(QCD_E_ARG, "cannot read state file extension");
        // n comes from the file: compare against what is left, so a huge
        // value cannot wrap the bound; a bad extension is treated as absent.
        if (memcmp(magic, kExtMagic, 8) == 0 && n <= size - end - 16) ext_bytes = n;
    }
    return QCD_OK;
}

}  // namespace

static int qcg_write(qcd_lattice* lat, const char* path, const qcd_run_info* info)
{
    if (!lat || !path || !info) return err(QCD_E_ARG, "null argument");
    qcd_params p;
    int rc = qcd_lattice_get_params(lat, &p);
    if (rc) return rc;
    unsigned counter = 0;
    if ((rc = qcd_prng_counter(lat, &counter))) return rc;
    const unsigned fsec = (info->get_fmunu || info->get_f0mu) ? 1u : 0u;
    const Layout lay = make_layout(p.sun, p.n, p.precision, info->iter, info->get_plaquettes, info->get_wilson,
                                   info->pl_level, fsec);
    const int N = p.sun;

    // lattice_make_bin_header (suncl.cpp:1939-1993)
    Header h = {};
    h.w[0] = str_word(kPrefix, 0);
    h.w[1] = str_word(kPrefix, 4);
    h.w[2] = str_word(kVersion, 0);
    h.w[4] = info->init;
    h.w[5] = (info->start == QCD_START_RANDOM) ? (uint32_t)QCD_START_HOT : info->start;   // model_start_hot
    h.w[6] = p.randseries;
    h.w[7] = generator_code(p.prng, p.ranlux_luxury);
    h.w[8] = rl_nskip(p.prng, p.ranlux_luxury);
    h.w[9] = counter;
    h.w[10] = info->nav;
    h.w[11] = info->nav_counter;
    h.w[12] = info->niter;
    h.w[13] = info->iter;
    h.w[14] = info->iter_counter;
    h.w[15] = (uint32_t)p.nhit;
    h.w[16] = info->wilson_r;
    h.w[17] = info->wilson_t;
    h.w[18] = (p.precision == QCD_FP32) ? 1u : 2u;   // convert_precision_to_uint (suncl.cpp:732-737)
    uint64_t bb;
    memcpy(&bb, &p.beta, 8);
    h.w[19] = (uint32_t)bb;                           // convert_to_uint_LOW / HIGH
    h.w[20] = (uint32_t)(bb >> 32);
    double phi = 0.0, omega = 0.0;
    if ((rc = qcd_lattice_get_twist(lat, &phi, &omega))) return rc;
    memcpy(&bb, &phi, 8);
    h.w[21] = (uint32_t)bb;
    h.w[22] = (uint32_t)(bb >> 32);
    memcpy(&bb, &omega, 8);
    h.w[23] = (uint32_t)bb;
    h.w[24] = (uint32_t)(bb >> 32);
    h.w[25] = info->pl_level;
    h.w[26] = info->fmunu_index1;
    h.w[27] = info->fmunu_index2;
    h.w[28] = info->get_fmunu ? 1u : 0u;
    h.w[29] = info->get_actions;
    h.w[30] = info->get_plaquettes;
    h.w[31] = info->get_wilson;
    h.w[32] = (uint32_t)lay.meas_size_F;
    h.w[33] = (uint32_t)lay.energies_size;
    h.w[34] = (uint32_t)lay.energies_size_F;
    h.w[35] = (uint32_t)lay.poly_size;
    h.w[36] = (uint32_t)N;
    h.w[37] = 4;
    for (int k = 0; k < 4; ++k) h.w[38 + k] = h.w[42 + k] = (uint32_t)p.n[k];

    // device data
    const size_t cap = (size_t)p.max_measurements;
    std::vector<double> plq(2 * cap), poly(4 * cap), wil(cap);
    int np = 0, nq = 0, nw = 0;
    if ((rc = qcd_read_measurements(lat, plq.data(), poly.data(), &np, &nq))) return rc;
    if ((rc = qcd_read_wilson(lat, wil.data(), &nw))) return rc;
    std::vector<double> fr(12 * cap);
    int nf = 0;
    if ((rc = qcd_read_fmunu(lat, fr.data(), &nf))) return rc;
    np = np < (int)cap ? np : (int)cap;
    nq = nq < (int)cap ? nq : (int)cap;
    nw = nw < (int)cap ? nw : (int)cap;
    std::vector<double> links(lay.V * 4 * (size_t)N * N * 2);
    if ((rc = qcd_lattice_export(lat, links.data()))) return rc;
    size_t img_bytes = 0;
    if ((rc = qcd_lattice_state_bytes(lat, &img_bytes))) return rc;
    std::vector<unsigned char> img(img_bytes);
    if ((rc = qcd_lattice_state_save(lat, img.data(), img_bytes))) return rc;

    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return err(QCD_E_ARG, "cannot write " + tmp + ": " + strerror(errno));
    bool ok = put(f, h.w, sizeof(h.w));
    {   // lattice_measurement: per-workgroup partials of the last reduction
        std::vector<double> z(2 * lay.meas_size_F, 0.0);
        ok = ok && put(f, z.data(), sizeof(double) * z.size());
    }
    const double V = (double)lay.V, bet = p.beta / N;
    {   // lattice_energies: action sums, bet * (3N - sum ReTr) per site and plane set
        std::vector<double> e(2 * lay.energies_size, 0.0);
        for (int i = 0; i < np && (uint64_t)i < lay.energies_size; ++i) {
            e[2 * i] = bet * (3.0 * N * V - plq[2 * i]);
            e[2 * i + 1] = bet * (3.0 * N * V - plq[2 * i + 1]);
        }
        ok = ok && put(f, e.data(), sizeof(double) * e.size());
    }
    if (lay.get_plq || lay.f) {   // lattice_energies_plq: (sum ReTr spatial, temporal), then F_munu
        std::vector<double> e(2 * lay.energies_size_F, 0.0);
        for (int i = 0; i < np && (uint64_t)i < lay.energies_size; ++i) {
            e[2 * i] = plq[2 * i];
            e[2 * i + 1] = plq[2 * i + 1];
            if (lay.f)
                for (int b = 0; b < 6; ++b)   // block 1 + 3 * generator + plane
                    for (int c = 0; c < 2; ++c) e[2 * ((1 + b) * lay.energies_size + i) + c] = fr[12 * i + 2 * b + c];
        }
        ok = ok && put(f, e.data(), sizeof(double) * e.size());
    }
    if (lay.get_wilson) {
        std::vector<double> e(lay.energies_size, 0.0);
        for (int i = 0; i < nw && (uint64_t)i < lay.energies_size; ++i) e[i] = wil[i];
        ok = ok && put(f, e.data(), sizeof(double) * e.size());
    }
    if (lay.pl_level > 0) {   // lattice_polyakov_loop, first half: (sum Re, sum Im)
        std::vector<double> e(2 * lay.poly_size, 0.0);
        for (int i = 0; i < nq && (uint64_t)i < lay.poly_size; ++i) {
            e[2 * i] = poly[4 * i];
            e[2 * i + 1] = poly[4 * i + 1];
        }
        ok = ok && put(f, e.data(), sizeof(double) * e.size());
    }
    {   // lattice_table: block j = 4 * plane + direction, ROWSIZE 4-vectors each
        const int planes = (N == 3) ? 3 : 1;
        std::vector<unsigned char> blk(lay.rowsize * lay.elem, 0);
        uint64_t written = 0;
        for (int j = 0; j < 4 * planes && ok; ++j) {
            const int k = j / 4, d = j % 4;
            std::fill(blk.begin(), blk.end(), 0);
            for (uint64_t s = 0; s < lay.V; ++s) {
                double pl[12];
                to_planes(N, &links[((s * 4 + d) * N * N) * 2], pl);
                if (lay.elem == 16) {
                    float* o = (float*)blk.data() + 4 * s;
                    for (int c = 0; c < 4; ++c) o[c] = (float)pl[4 * k + c];
                } else {
                    double* o = (double*)blk.data() + 4 * s;
                    for (int c = 0; c < 4; ++c) o[c] = pl[4 * k + c];
                }
            }
            ok = ok && put(f, blk.data(), blk.size());
            written += lay.rowsize;
        }
        std::fill(blk.begin(), blk.end(), 0);
        while (ok && written < lay.table_size) {
            const uint64_t n = std::min<uint64_t>(lay.table_size - written, lay.rowsize);
            ok = put(f, blk.data(), n * lay.elem);
            written += n;
        }
    }
    {
        const uint64_t n = img.size();
        ok = ok && put(f, kExtMagic, 8) && put(f, &n, 8) && put(f, img.data(), img.size());
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        remove(tmp.c_str());
        return err(QCD_E_ARG, "writing " + tmp + " failed");
    }
    if (rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return err(QCD_E_ARG, std::string("cannot rename into ") + path + ": " + strerror(errno));
    }
    return QCD_OK;
}

static int qcg_read_header(const char* path, qcd_params* params, qcd_run_info* info, int* has_exact)
{
    File fh;
    Header h;
    qcd_params q;
    Layout lay;
    uint64_t ext = 0;
    int rc = open_state(path, fh, h, q, info, lay, ext);
    if (rc) return rc;
    if (ext) {
        unsigned char head[136];   // the image head, parameter fields included
        qcd_params e = q;
        const size_t n = ext < sizeof(head) ? (size_t)ext : sizeof(head);
        if (!get(fh.f, head, n) || !image_params(head, n, &e)) return err(QCD_E_ARG, "bad state image head");
        q = e;
    }
    if (params) *params = q;
    if (has_exact) *has_exact = ext ? 1 : 0;
    return QCD_OK;
}

static int qcg_read(qcd_lattice* lat, const char* path, qcd_run_info* info)
{
    if (!lat) return err(QCD_E_ARG, "null lattice");
    File fh;
    Header h;
    qcd_params q;
    Layout lay;
    uint64_t ext = 0;
    qcd_run_info ri = {};
    int rc = open_state(path, fh, h, q, &ri, lay, ext);
    if (rc) return rc;
    qcd_params lp;
    if ((rc = qcd_lattice_get_params(lat, &lp))) return rc;
    // the quantities the header fixes must match the lattice
    bool fits = lp.sun == q.sun && lp.precision == q.precision && lp.prng == q.prng && lp.nhit == q.nhit &&
                lp.randseries == q.randseries && memcmp(&lp.beta, &q.beta, 8) == 0;
    if (lp.prng == QCD_PRNG_RANLUX) fits = fits && lp.ranlux_luxury == q.ranlux_luxury;
    for (int k = 0; k < 4; ++k) fits = fits && lp.n[k] == q.n[k];
    if (!fits) return err(QCD_E_ARG, std::string(path) + " holds another lattice (group, sizes, precision, beta, "
                                                         "NHIT or generator differ)");
    // A hot (ints = 0) run is the reference's random start unless the file's
    // own image says it began with the Haar start (cold-size productions).
    int start = (int)ri.start;
    if (ext) {
        std::vector<unsigned char> img((size_t)ext);
        qcd_params e;
        if (!get(fh.f, img.data(), img.size())) return err(QCD_E_ARG, "short state image");
        uint32_t flags = 0;
        if (start == QCD_START_HOT && !(image_start_flags(img.data(), img.size(), &flags) && !(flags & 2u)))
            start = QCD_START_RANDOM;
        if (image_params(img.data(), img.size(), &e) && same_params(e, lp)) {
            if ((rc = qcd_lattice_state_load(lat, img.data(), img.size()))) return rc;
            if ((rc = apply_settings(lat, ri))) return rc;
            if (info) *info = ri;
            return QCD_OK;
        }
    } else if (start == QCD_START_HOT) {
        start = QCD_START_RANDOM;   // a reference file
    }
    // reference sections (lattice_load_state, suncl.cpp:2060-2072)
    const int N = q.sun;
    if (fseeko(fh.f, (off_t)sizeof(h.w), SEEK_SET) != 0) return err(QCD_E_ARG, "cannot seek in state file");
    std::vector<double> skip(2 * lay.meas_size_F);
    bool ok = get(fh.f, skip.data(), sizeof(double) * skip.size());
    std::vector<double> en(2 * lay.energies_size);
    ok = ok && get(fh.f, en.data(), sizeof(double) * en.size());   // action sums: derived, not kept
    std::vector<double> eplq, ewil, epoly;
    if (lay.get_plq || lay.f) {
        eplq.resize(2 * lay.energies_size_F);
        ok = ok && get(fh.f, eplq.data(), sizeof(double) * eplq.size());
    }
    if (lay.get_wilson) {
        ewil.resize(lay.energies_size);
        ok = ok && get(fh.f, ewil.data(), sizeof(double) * ewil.size());
    }
    if (lay.pl_level > 0) {
        epoly.resize(2 * lay.poly_size);
        ok = ok && get(fh.f, epoly.data(), sizeof(double) * epoly.size());
    }
    if (!ok) return err(QCD_E_ARG, "short measurement sections");
    std::vector<double> links(lay.V * 4 * (size_t)N * N * 2);
    {
        const int planes = (N == 3) ? 3 : 1;
        std::vector<double> pl(lay.V * 4 * 4 * (size_t)planes);   // [site][dir][plane*4 + c]
        std::vector<unsigned char> blk(lay.rowsize * lay.elem);
        for (int j = 0; j < 4 * planes; ++j) {
            if (!get(fh.f, blk.data(), blk.size())) return err(QCD_E_ARG, "short table section");
            const int k = j / 4, d = j % 4;
            for (uint64_t s = 0; s < lay.V; ++s)
                for (int c = 0; c < 4; ++c)
                    pl[(s * 4 + d) * 4 * planes + 4 * k + c] =
                        (lay.elem == 16) ? (double)((const float*)blk.data())[4 * s + c]
                                         : ((const double*)blk.data())[4 * s + c];
        }
        for (uint64_t s = 0; s < lay.V; ++s)
            for (int d = 0; d < 4; ++d)
                from_planes(N, &pl[(s * 4 + d) * 4 * planes], &links[((s * 4 + d) * N * N) * 2]);
    }
    if ((rc = qcd_lattice_init(lat, start == QCD_START_GID ? QCD_START_GID : QCD_START_COLD))) return rc;
    if ((rc = qcd_lattice_import(lat, links.data()))) return rc;
    const int n = (int)ri.iter_counter;
    const int np = (lay.get_plq || lay.f) ? (int)std::min<uint64_t>((uint64_t)n, lay.energies_size) : 0;
    const int nw = lay.get_wilson ? (int)std::min<uint64_t>((uint64_t)n, lay.energies_size) : 0;
    const int nq = lay.pl_level > 0 ? (int)std::min<uint64_t>((uint64_t)n, lay.poly_size) : 0;
    std::vector<double> poly4(4 * (size_t)(nq ? nq : 1), 0.0);
    for (int i = 0; i < nq; ++i) {
        poly4[4 * i] = epoly[2 * i];
        poly4[4 * i + 1] = epoly[2 * i + 1];
    }
    if ((rc = qcd_write_measurements(lat, np ? eplq.data() : nullptr, nq ? poly4.data() : nullptr,
                                     nw ? ewil.data() : nullptr, np, nq, nw)))
        return rc;
    if (lay.f && np) {
        std::vector<double> fr(12 * (size_t)np);
        for (int i = 0; i < np; ++i)
            for (int b = 0; b < 6; ++b)
                for (int c = 0; c < 2; ++c) fr[12 * i + 2 * b + c] = eplq[2 * ((1 + b) * lay.energies_size + i) + c];
        if ((rc = qcdhip::write_fmunu(lat, fr.data(), np))) return rc;
    }
    if ((rc = apply_settings(lat, ri))) return rc;
    if (ri.wilson_r >= 1 && ri.wilson_t >= 1 &&
        (rc = qcd_lattice_set_wilson(lat, (int)ri.wilson_r, (int)ri.wilson_t)))
        return rc;
    if (start != QCD_START_GID && (rc = qcd_prng_replay(lat, ri.prng_counter, start))) return rc;
    if (info) *info = ri;
    return QCD_OK;
}

// The C ABI: no C++ exception (a failed allocation for a corrupt size, a
// stream error) leaves these entry points; it becomes an error code.
extern "C" {

int qcd_qcg_write(qcd_lattice* lat, const char* path, const qcd_run_info* info)
{
    try {
        return qcg_write(lat, path, info);
    } catch (const std::exception& e) {
        return err(QCD_E_NOMEM, std::string("qcd_qcg_write: ") + e.what());
    }
}

int qcd_qcg_read_header(const char* path, qcd_params* params, qcd_run_info* info, int* has_exact)
{
    try {
        return qcg_read_header(path, params, info, has_exact);
    } catch (const std::exception& e) {
        return err(QCD_E_ARG, std::string("qcd_qcg_read_header: ") + e.what());
    }
}

int qcd_qcg_read(qcd_lattice* lat, const char* path, qcd_run_info* info)
{
    try {
        return qcg_read(lat, path, info);
    } catch (const std::exception& e) {
        return err(QCD_E_ARG, std::string("qcd_qcg_read: ") + e.what());
    }
}

}  // extern "C"
