// qcd_flow.hip -- learned global proposal for SU(2) lattices: a
// gauge-equivariant normalising flow of masked coupling layers, evaluated
// with BF16 MFMA convolutions (FP32 accumulation), for an independence
// Metropolis step with the exact Wilson action.  No counterpart in the
// reference (new work); its plain heatbath is the fallback and the judge.
//
// Links: unit quaternions q = (a0, a1, a2, a3), U = a0 + i (a1 s1 + a2 s2 +
// a3 s3) = [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]], stored
// links[(b * 4 + mu) * V + site] (float4), site = y + z N2 + t N2 N3 +
// x N2 N3 N4 (the canonical order of the C ABI).
//
// Coupling layer l (direction mu = l % 4, mask parity m = (l / 4) % 2)
// updates U_mu(x) at the active sites, sum_{nu != mu} x_nu + m even.  The
// staple sum A(x) = sum_nu [U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ +
// U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)] holds frozen links only (the
// U_mu(x +- nu) sit on inactive sites), A = k V with V in SU(2), and the
// local action is -beta k (U V)_0.  W = U V = exp(i theta n.s) is gauge
// covariant; the layer maps theta in [0, pi] by a monotone rational-quadratic
// spline f whose parameters depend on frozen, gauge-invariant features, and
// sets U' = exp(i f(theta) n.s) V^+.  Haar measure in (theta, n) is
// sin^2(theta) d(theta) d(Omega), so log J = log f'(theta) + 2 log(sin
// f(theta) / sin theta).  log q(U) = -sum_layers log J (Haar prior, density
// 1).
//
// Conditioner per layer: features per site (channels-last, BF16, padded to
// 8): k(x)/6 at active sites (0 elsewhere), (P_nu rho)_0 of the three
// plaquettes in the planes orthogonal to mu, the active mask; then a 3^4
// stencil convolution to 32 channels + bias, tanh; then a 3^4 stencil
// convolution to 16 channels + bias (3K + 1 spline parameters used).
// Stencil point st = ((ox * 3 + oy) * 3 + oz) * 3 + ot, offsets o - 1 along
// x, y, z, t, periodic.  Convolutions are implicit GEMMs on
// v_mfma_f32_16x16x16_bf16: rows = sites, K = (stencil point, input
// channel), N = output channels.
//
// Spline (Durkan et al. 2019, rational-quadratic, K bins on [0, pi]): raw
// outputs r[0..K) widths, r[K..2K) heights, r[2K..3K] derivatives; widths =
// pi (eps + (1 - K eps) softmax), heights likewise, derivatives = eps_d +
// (1 - eps_d) softplus(r) / ln 2 (all-zero outputs give the identity map).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qcd_hip.h"
#include "qcd_internal.h"

namespace {

constexpr int kStencil = 81;
constexpr int kCin0 = 8;     // input feature channels (5 used)
constexpr int kHidden = 32;
constexpr int kOut = 16;     // spline parameter channels (3K + 1 used)
constexpr int kMaxBins = 5;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kMinW = 1e-3f, kMinD = 1e-3f;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct FGeom {
    int n1, n2, n3, n4;   // extents along x, y, z, t
    uint32_t V;
};

__host__ __device__ inline uint16_t f2bf(float f)
{   // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline void coords(const FGeom& g, uint32_t s, int c[4])
{
    c[1] = (int)(s % (uint32_t)g.n2);
    uint32_t r = s / (uint32_t)g.n2;
    c[2] = (int)(r % (uint32_t)g.n3);
    r /= (uint32_t)g.n3;
    c[3] = (int)(r % (uint32_t)g.n4);
    c[0] = (int)(r / (uint32_t)g.n4);
}
__device__ inline uint32_t site(const FGeom& g, const int c[4])
{
    return (uint32_t)c[1] + (uint32_t)g.n2 * ((uint32_t)c[2] + (uint32_t)g.n3 * ((uint32_t)c[3] + (uint32_t)g.n4 * (uint32_t)c[0]));
}
__device__ inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }
__device__ inline uint32_t shifted(const FGeom& g, const int c[4], int d, int o)
{
    int e[4] = {c[0], c[1], c[2], c[3]};
    const int n[4] = {g.n1, g.n2, g.n3, g.n4};
    e[d] = wrap(e[d] + o, n[d]);
    return site(g, e);
}

// quaternion algebra for U = a0 + i a.s
__device__ inline float4 qmul(float4 a, float4 b)
{
    return make_float4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w,
                       a.x * b.y + b.x * a.y - (a.z * b.w - a.w * b.z),
                       a.x * b.z + b.x * a.z - (a.w * b.y - a.y * b.w),
                       a.x * b.w + b.x * a.w - (a.y * b.z - a.z * b.y));
}
__device__ inline float4 qdag(float4 a) { return make_float4(a.x, -a.y, -a.z, -a.w); }
__device__ inline float4 qadd(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

__device__ inline float4 link(const float4* L, const FGeom& g, uint32_t b, int mu, uint32_t s)
{
    return L[((size_t)b * 4 + mu) * g.V + s];
}

// staple sum of U_mu(x) (see the file comment)
__device__ float4 staple(const float4* L, const FGeom& g, uint32_t b, int mu, const int c[4])
{
    float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t x = site(g, c);
    const uint32_t xp = shifted(g, c, mu, 1);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        const uint32_t xpn = shifted(g, c, nu, 1), xmn = shifted(g, c, nu, -1);
        int cm[4] = {c[0], c[1], c[2], c[3]};
        const int n[4] = {g.n1, g.n2, g.n3, g.n4};
        cm[nu] = wrap(cm[nu] - 1, n[nu]);
        const uint32_t xpmn = shifted(g, cm, mu, 1);   // x + mu - nu
        const float4 f = qmul(qmul(link(L, g, b, nu, xp), qdag(link(L, g, b, mu, xpn))), qdag(link(L, g, b, nu, x)));
        const float4 r = qmul(qmul(qdag(link(L, g, b, nu, xpmn)), qdag(link(L, g, b, mu, xmn))), link(L, g, b, nu, xmn));
        A = qadd(A, qadd(f, r));
    }
    return A;
}

// (U_nu(x) U_rho(x+nu) U_nu(x+rho)^+ U_rho(x)^+)_0
__device__ float plaq0(const float4* L, const FGeom& g, uint32_t b, int nu, int rho, const int c[4])
{
    const uint32_t x = site(g, c);
    const float4 p = qmul(qmul(link(L, g, b, nu, x), link(L, g, b, rho, shifted(g, c, nu, 1))),
                          qmul(qdag(link(L, g, b, nu, shifted(g, c, rho, 1))), qdag(link(L, g, b, rho, x))));
    return p.x;
}

__device__ inline bool is_active(const int c[4], int mu, int par)
{
    int s = par;
    for (int d = 0; d < 4; ++d)
        if (d != mu) s += c[d];
    return (s & 1) == 0;
}

// ---------------------------------------------------------------- kernels
// splitmix64's output function (a bijection of 64-bit words).
__host__ __device__ inline uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// 2 * n uniforms in (0, 1) (24 bits each) for item `j` of the draw keyed by
// `key`: counter-based, the counter (key, j) hashed before the splitmix64
// sequence starts, so neither neighbouring items nor draws with
// neighbouring keys share numbers.
template <int n>
__device__ inline void uniforms(uint64_t key, uint64_t j, float* u)
{
    uint64_t z = mix64(key ^ mix64(j + 0x632BE59BD9B4E019ull));
    for (int k = 0; k < n; ++k) {
        z += 0x9E3779B97F4A7C15ull;
        const uint64_t x = mix64(z);
        u[2 * k] = ((float)(uint32_t)(x >> 40) + 0.5f) * 5.9604644775390625e-8f;
        u[2 * k + 1] = ((float)(uint32_t)((x >> 8) & 0xffffffu) + 0.5f) * 5.9604644775390625e-8f;
    }
}

__global__ void k_flow_haar(float4* L, uint64_t n, uint64_t seed)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float u[4];
    uniforms<2>(mix64(seed), j, u);
    // Box-Muller: four Gaussians, normalised: uniform on S^3 = Haar on SU(2)
    const float r1 = sqrtf(-2.f * logf(u[0])), r2 = sqrtf(-2.f * logf(u[2]));
    float s1, c1, s2, c2;
    sincosf(2.f * kPi * u[1], &s1, &c1);
    sincosf(2.f * kPi * u[3], &s2, &c2);
    float4 q = make_float4(r1 * c1, r1 * s1, r2 * c2, r2 * s2);
    const float nrm = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    L[j] = make_float4(q.x * nrm, q.y * nrm, q.z * nrm, q.w * nrm);
}

__global__ void k_flow_features(const float4* L, uint16_t* feat, FGeom g, uint32_t B, int mu, int par)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * g.V) return;
    const uint32_t b = j / g.V, s = j % g.V;
    int c[4];
    coords(g, s, c);
    const bool act = is_active(c, mu, par);
    float f[kCin0] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (act) {
        const float4 A = staple(L, g, b, mu, c);
        f[0] = sqrtf(A.x * A.x + A.y * A.y + A.z * A.z + A.w * A.w) * (1.f / 6.f);
        f[4] = 1.f;
    }
    int k = 1;
    for (int nu = 0; nu < 4; ++nu)
        for (int rho = nu + 1; rho < 4; ++rho)
            if (nu != mu && rho != mu) f[k++] = plaq0(L, g, b, nu, rho, c);
    uint16_t* o = feat + (size_t)j * kCin0;
    for (int q = 0; q < kCin0; ++q) o[q] = f2bf(f[q]);
}

// One wavefront: a tile of 16 sites x 16 output channels.  in[(row) * cin +
// channel] (BF16), wt[n * (81 cin) + k] (BF16, k = st * cin + channel),
// out: tanh -> BF16 [row * cout + n] (act) or FP32 [row * cout + n].
template <int CIN>
__global__ __launch_bounds__(256) void k_flow_conv(const uint16_t* in, int cin, const uint16_t* wt, const float* bias,
                                                   int cout, uint16_t* out_bf, float* out_f, FGeom g, uint32_t B,
                                                   int act)
{
    static_assert(CIN % 4 == 0, "a lane's 4 values stay inside one stencil point");
    (void)cin;   // == CIN (the host dispatches on it)
    const uint32_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int lane = threadIdx.x & 63;
    const uint32_t ntile_m = B * g.V / 16, ntile_n = (uint32_t)cout / 16;
    if (wave >= ntile_m * ntile_n) return;   // whole wavefront
    const uint32_t tm = wave % ntile_m, tn = wave / ntile_m;
    const uint32_t row = tm * 16 + (lane & 15);   // A row of this lane
    const uint32_t b = row / g.V, s = row % g.V;
    int c[4];
    coords(g, s, c);
    const int n[4] = {g.n1, g.n2, g.n3, g.n4};
    const int kq = 4 * (lane >> 4);                // k offset of this lane's 4 values
    const uint32_t ncol = tn * 16 + (lane & 15);   // B column of this lane
    constexpr int K = kStencil * CIN;
    const uint16_t* wrow = wt + (size_t)ncol * K;
    const uint16_t* inb = in + (size_t)b * g.V * CIN;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    // The operands of U consecutive k-steps are loaded before their MFMAs,
    // so U gathers are in flight at once; the MFMAs run in k order, as one
    // step at a time would.
    constexpr int U = 8;
    for (int k0 = 0; k0 < K; k0 += 16 * U) {
        s4 av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 16 * u + kq;   // K is a multiple of 4: a lane's 4 values are all inside or all past K
            av[u] = s4{0, 0, 0, 0};
            bv[u] = s4{0, 0, 0, 0};
            if (k < K) {
                const int st = k / CIN, ch = k % CIN;
                const int o[4] = {st / 27 - 1, (st / 9) % 3 - 1, (st / 3) % 3 - 1, st % 3 - 1};
                int e[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) e[d] = wrap(c[d] + o[d], n[d]);
                av[u] = *reinterpret_cast<const s4*>(inb + (size_t)site(g, e) * CIN + ch);
                bv[u] = *reinterpret_cast<const s4*>(wrow + k);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[u], bv[u], acc, 0, 0, 0);
    }
    // lane holds D[4 * (lane / 16) + i][lane % 16]
    for (int i = 0; i < 4; ++i) {
        const uint32_t r = tm * 16 + 4 * (lane >> 4) + i;
        float v = acc[i] + bias[ncol];
        if (act) out_bf[(size_t)r * cout + ncol] = f2bf(tanhf(v));
        else out_f[(size_t)r * cout + ncol] = v;
    }
}

struct Spline {
    float xk[kMaxBins + 1], yk[kMaxBins + 1], d[kMaxBins + 1];
    int K;
};

__device__ inline float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }

__device__ void make_spline(const float* r, int K, Spline& sp)
{
    sp.K = K;
    float mw = r[0], mh = r[K];
    for (int i = 1; i < K; ++i) {
        mw = fmaxf(mw, r[i]);
        mh = fmaxf(mh, r[K + i]);
    }
    float ew[kMaxBins], eh[kMaxBins], sw = 0.f, sh = 0.f;
    for (int i = 0; i < K; ++i) {
        ew[i] = expf(r[i] - mw);
        eh[i] = expf(r[K + i] - mh);
        sw += ew[i];
        sh += eh[i];
    }
    sp.xk[0] = sp.yk[0] = 0.f;
    for (int i = 0; i < K; ++i) {
        sp.xk[i + 1] = sp.xk[i] + kPi * (kMinW + (1.f - K * kMinW) * ew[i] / sw);
        sp.yk[i + 1] = sp.yk[i] + kPi * (kMinW + (1.f - K * kMinW) * eh[i] / sh);
    }
    sp.xk[K] = sp.yk[K] = kPi;
    for (int i = 0; i <= K; ++i) sp.d[i] = kMinD + (1.f - kMinD) * softplus(r[2 * K + i]) * 1.4426950408889634f;
}

// forward: y = f(x), dlog = log f'(x)
__device__ float spline_fwd(const Spline& sp, float x, float& dlog)
{
    int k = 0;
    while (k < sp.K - 1 && x >= sp.xk[k + 1]) ++k;
    const float w = sp.xk[k + 1] - sp.xk[k], h = sp.yk[k + 1] - sp.yk[k], s = h / w;
    const float xi = fminf(fmaxf((x - sp.xk[k]) / w, 0.f), 1.f), om = 1.f - xi;
    const float den = s + (sp.d[k + 1] + sp.d[k] - 2.f * s) * xi * om;
    const float y = sp.yk[k] + h * (s * xi * xi + sp.d[k] * xi * om) / den;
    const float num = s * s * (sp.d[k + 1] * xi * xi + 2.f * s * xi * om + sp.d[k] * om * om);
    dlog = logf(num) - 2.f * logf(den);
    return y;
}

// inverse: x = f^-1(y), dlog = log f'(x)
__device__ float spline_inv(const Spline& sp, float y, float& dlog)
{
    int k = 0;
    while (k < sp.K - 1 && y >= sp.yk[k + 1]) ++k;
    const float w = sp.xk[k + 1] - sp.xk[k], h = sp.yk[k + 1] - sp.yk[k], s = h / w;
    const float dy = y - sp.yk[k], dd = sp.d[k + 1] + sp.d[k] - 2.f * s;
    const float a = h * (s - sp.d[k]) + dy * dd;
    const float bq = h * sp.d[k] - dy * dd;
    const float cq = -s * dy;
    const float disc = fmaxf(bq * bq - 4.f * a * cq, 0.f);
    const float xi = fminf(fmaxf(2.f * cq / (-bq - sqrtf(disc)), 0.f), 1.f), om = 1.f - xi;
    const float den = s + dd * xi * om;
    const float num = s * s * (sp.d[k + 1] * xi * xi + 2.f * s * xi * om + sp.d[k] * om * om);
    dlog = logf(num) - 2.f * logf(den);
    return sp.xk[k] + xi * w;
}

// Coupling of layer (mu, par): forward (inverse = 0) maps theta -> f(theta),
// inverse maps theta' -> f^-1(theta'); both add log J = log f'(theta) +
// 2 log(sin f(theta) / sin theta) of the forward map to logj[row].
__global__ void k_flow_couple(float4* L, const float* prm, float* logj, FGeom g, uint32_t B, int mu, int par, int K,
                              int inverse)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * g.V) return;
    const uint32_t b = j / g.V, s = j % g.V;
    int c[4];
    coords(g, s, c);
    if (!is_active(c, mu, par)) return;
    const float4 A = staple(L, g, b, mu, c);
    const float k = sqrtf(A.x * A.x + A.y * A.y + A.z * A.z + A.w * A.w);
    if (!(k > 1e-20f)) return;
    const float ik = 1.f / k;
    const float4 Vq = make_float4(A.x * ik, A.y * ik, A.z * ik, A.w * ik);
    float4& U = L[((size_t)b * 4 + mu) * g.V + s];
    const float4 W = qmul(U, Vq);
    const float sn = sqrtf(W.y * W.y + W.z * W.z + W.w * W.w);
    const float th = atan2f(sn, W.x);
    Spline sp;
    make_spline(prm + (size_t)j * kOut, K, sp);
    float dlog, th2;
    if (!inverse) th2 = spline_fwd(sp, th, dlog);
    else th2 = spline_inv(sp, th, dlog);
    const float th_in = inverse ? th2 : th, th_out = inverse ? th : th2;   // theta and f(theta)
    const float si = sinf(th_in), so = sinf(th_out);
    const float ratio = (si > 1e-6f && so > 1e-6f) ? so / si : expf(dlog);   // limit at the poles: f'(theta)
    logj[j] += dlog + 2.f * logf(ratio);
    float4 Wn;
    if (sn > 1e-12f) {
        const float f = sinf(th2) / sn;
        Wn = make_float4(cosf(th2), W.y * f, W.z * f, W.w * f);
    } else {
        Wn = make_float4(cosf(th2) >= 0.f ? 1.f : -1.f, 0.f, 0.f, 0.f);
    }
    U = qmul(Wn, qdag(Vq));
}

// log J = log f'(theta) + 2 log(sin f(theta) / sin theta) of the forward map
// at theta (theta_in) with image theta_out; at the poles the ratio's limit
// f'(theta).
__device__ inline float log_jac(float dlog, float th_in, float th_out)
{
    const float si = sinf(th_in), so = sinf(th_out);
    const float ratio = (si > 1e-6f && so > 1e-6f) ? so / si : expf(dlog);
    return dlog + 2.f * logf(ratio);
}

// Learned Metropolis update of layer (mu, par)'s active links of
// configuration 0 (a Metropolis-within-Gibbs step on each).  The layer's
// spline parameters depend on frozen links only, and no plaquette holds two
// active links, so given the frozen links the target factorises into
// exp(beta k (U V)_0) per active link and the layer's map, applied to a
// Haar draw, proposes each active link independently with density
// q_a(U) = exp(-log J(f^-1(theta(U V)))) (w.r.t. Haar).  Each active link is
// replaced by its proposal U' with probability
// min(1, exp(beta k ((U'V)_0 - (UV)_0)) q_a(U) / q_a(U')): detailed balance
// for every link, so the Wilson-action distribution stays exact.  Counts
// accepted links in *accepted (one atomic per wavefront).
// The sweep's seed: mode 0 sets it, 1 advances it by one and counts a
// sweep (a recorded sweep), 2 sets it and counts a sweep.
__global__ void k_flow_imh_key(unsigned long long* imh, unsigned long long seed, int mode)
{
    if (threadIdx.x == 0) {
        imh[2] = mode == 1 ? imh[2] + 1ull : seed;
        if (mode) imh[1] += 1ull;
    }
}

__global__ __launch_bounds__(256) void k_flow_imh_couple(float4* L, const float* prm, FGeom g, int mu, int par, int K,
                                                         float beta, uint64_t lmix, unsigned long long* imh)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    // the layer's draw key: (sweep seed, layer) hashed
    const uint64_t seed = mix64(mix64((uint64_t)imh[2]) ^ lmix);
    unsigned long long* accepted = imh;
    bool acc = false;
    if (s < g.V) {
        int c[4];
        coords(g, s, c);
        if (is_active(c, mu, par)) {
            const float4 A = staple(L, g, 0, mu, c);
            const float k = sqrtf(A.x * A.x + A.y * A.y + A.z * A.z + A.w * A.w);
            if (k > 1e-20f) {
                const float ik = 1.f / k;
                const float4 Vq = make_float4(A.x * ik, A.y * ik, A.z * ik, A.w * ik);
                float4& U = L[(size_t)mu * g.V + s];
                const float4 W = qmul(U, Vq);
                const float th = atan2f(sqrtf(W.y * W.y + W.z * W.z + W.w * W.w), W.x);
                Spline sp;
                make_spline(prm + (size_t)s * kOut, K, sp);
                float dlog;
                const float thz = spline_inv(sp, th, dlog);
                const float lj_cur = log_jac(dlog, thz, th);
                // the proposal: a Haar element z = (cos thz', sin thz' n') in the W frame, mapped
                float u[6];
                uniforms<3>(seed, s, u);
                const float r1 = sqrtf(-2.f * logf(u[0])), r2 = sqrtf(-2.f * logf(u[2]));
                float s1, c1, s2, c2;
                sincosf(2.f * kPi * u[1], &s1, &c1);
                sincosf(2.f * kPi * u[3], &s2, &c2);
                const float4 zq = make_float4(r1 * c1, r1 * s1, r2 * c2, r2 * s2);
                const float zs = sqrtf(zq.y * zq.y + zq.z * zq.z + zq.w * zq.w);
                const float thz2 = atan2f(zs, zq.x);
                float dlog2;
                const float th2 = spline_fwd(sp, thz2, dlog2);
                const float lj_prop = log_jac(dlog2, thz2, th2);
                // log a = beta k (cos th2 - cos th) + log q(cur) - log q(prop)
                const float log_a = beta * k * (cosf(th2) - cosf(th)) - lj_cur + lj_prop;
                if (logf(u[4]) < log_a) {
                    float4 Wn;
                    if (zs > 1e-12f) {
                        const float f = sinf(th2) / zs;
                        Wn = make_float4(cosf(th2), zq.y * f, zq.z * f, zq.w * f);
                    } else {
                        Wn = make_float4(cosf(th2) >= 0.f ? 1.f : -1.f, 0.f, 0.f, 0.f);
                    }
                    const float4 Un = qmul(Wn, qdag(Vq));
                    const float nn = rsqrtf(Un.x * Un.x + Un.y * Un.y + Un.z * Un.z + Un.w * Un.w);
                    U = make_float4(Un.x * nn, Un.y * nn, Un.z * nn, Un.w * nn);
                    acc = true;
                }
            }
        }
    }
    const unsigned long long m = __ballot(acc);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(accepted, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------- training
// Layer-local forward KL: for every coupling layer l, the mean over data
// configurations of sum over active links of log J(f^-1(theta)) =
// -log q_a(U_a | frozen links), the density with which the learned
// Metropolis sweep proposes each active link.  The layer's input is the data
// configuration itself, so a layer's loss depends on its own parameters
// only, and the gradient flows back through conv2, tanh and conv1 alone.

// -log q_a of one active link in double precision (the map of
// make_spline / spline_inv / log_jac, written for double): theta = angle of
// U V, r = the 3K + 1 raw spline parameters.
__device__ double nll_link(const double* r, int K, double th)
{
    const double pi = 3.14159265358979323846;
    double mw = r[0], mh = r[K];
    for (int i = 1; i < K; ++i) {
        mw = fmax(mw, r[i]);
        mh = fmax(mh, r[K + i]);
    }
    double ew[kMaxBins], eh[kMaxBins], sw = 0.0, sh = 0.0;
    for (int i = 0; i < K; ++i) {
        ew[i] = exp(r[i] - mw);
        eh[i] = exp(r[K + i] - mh);
        sw += ew[i];
        sh += eh[i];
    }
    double xk[kMaxBins + 1], yk[kMaxBins + 1], d[kMaxBins + 1];
    xk[0] = yk[0] = 0.0;
    for (int i = 0; i < K; ++i) {
        xk[i + 1] = xk[i] + pi * (kMinW + (1.0 - K * kMinW) * ew[i] / sw);
        yk[i + 1] = yk[i] + pi * (kMinW + (1.0 - K * kMinW) * eh[i] / sh);
    }
    xk[K] = yk[K] = pi;
    for (int i = 0; i <= K; ++i) {
        const double x = r[2 * K + i];
        d[i] = kMinD + (1.0 - kMinD) * (x > 30.0 ? x : log1p(exp(x))) * 1.4426950408889634;
    }
    int k = 0;
    while (k < K - 1 && th >= yk[k + 1]) ++k;
    const double w = xk[k + 1] - xk[k], h = yk[k + 1] - yk[k], sl = h / w;
    const double dy = th - yk[k], dd = d[k + 1] + d[k] - 2.0 * sl;
    const double a = h * (sl - d[k]) + dy * dd, bq = h * d[k] - dy * dd, cq = -sl * dy;
    const double disc = fmax(bq * bq - 4.0 * a * cq, 0.0);
    const double xi = fmin(fmax(2.0 * cq / (-bq - sqrt(disc)), 0.0), 1.0), om = 1.0 - xi;
    const double den = sl + dd * xi * om;
    const double num = sl * sl * (d[k + 1] * xi * xi + 2.0 * sl * xi * om + d[k] * om * om);
    const double dlog = log(num) - 2.0 * log(den);
    const double thz = xk[k] + xi * w;
    const double si = sin(thz), so = sin(th);
    const double ratio = (si > 1e-9 && so > 1e-9) ? so / si : exp(dlog);
    return dlog + 2.0 * log(ratio);
}

// Per site of a batch: at active sites, -log q_a and its gradient with
// respect to the raw spline parameters (central differences in double
// precision, step 1e-5; the map is smooth except at bin edges, a set of
// measure zero) into g (BF16, [row][16], zero at inactive sites and for
// the unused channels); loss[row] = -log q_a (0 at inactive sites).
__global__ void k_flow_nll_grad(const float4* L, const float* prm, uint16_t* g, float* loss, FGeom g4, uint32_t B,
                                int mu, int par, int K)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * g4.V) return;
    const uint32_t b = j / g4.V, s = j % g4.V;
    int c[4];
    coords(g4, s, c);
    float gr[kOut];
    for (int q = 0; q < kOut; ++q) gr[q] = 0.f;
    float lo = 0.f;
    if (is_active(c, mu, par)) {
        const float4 A = staple(L, g4, b, mu, c);
        const float k = sqrtf(A.x * A.x + A.y * A.y + A.z * A.z + A.w * A.w);
        if (k > 1e-20f) {
            const float ik = 1.f / k;
            const float4 W = qmul(L[((size_t)b * 4 + mu) * g4.V + s], make_float4(A.x * ik, A.y * ik, A.z * ik, A.w * ik));
            const double th = atan2((double)sqrtf(W.y * W.y + W.z * W.z + W.w * W.w), (double)W.x);
            double r[3 * kMaxBins + 1];
            const int np = 3 * K + 1;
            for (int q = 0; q < np; ++q) r[q] = (double)prm[(size_t)j * kOut + q];
            lo = (float)nll_link(r, K, th);
            const double hstep = 1e-5;
            for (int q = 0; q < np; ++q) {
                const double r0 = r[q];
                r[q] = r0 + hstep;
                const double up = nll_link(r, K, th);
                r[q] = r0 - hstep;
                const double dn = nll_link(r, K, th);
                r[q] = r0;
                gr[q] = (float)((up - dn) / (2.0 * hstep));
            }
        }
    }
    loss[j] = lo;
    for (int q = 0; q < kOut; ++q) g[(size_t)j * kOut + q] = f2bf(gr[q]);
}

// Weight gradient dW[n][k] = sum over rows r of gout[r][n] X[r][k], X the
// stencil (im2col) view of `in` (X[r][st * cin + ch] = in[nbr(r, st)][ch]),
// as v_mfma_f32_16x16x16_bf16 tiles: D[n][k] += A[n][r] B[r][k] with the
// rows r as the reduction dimension.  One wavefront per (16 x 16 tile,
// chunk of `chunk` rows); partial sums to part[chunk_index][n][k].
__global__ __launch_bounds__(256) void k_flow_wgrad(const uint16_t* gout, int cout, const uint16_t* in, int cin,
                                                    float* part, FGeom g, uint32_t B, uint32_t chunk)
{
    const uint32_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int lane = threadIdx.x & 63;
    const int K = kStencil * cin;
    const uint32_t tn = (uint32_t)(K + 15) / 16, tm = (uint32_t)cout / 16;
    const uint32_t R = B * g.V, nchunk = (R + chunk - 1) / chunk;
    if (wave >= tn * tm * nchunk) return;   // whole wavefront
    const uint32_t ch_i = wave / (tn * tm), t = wave % (tn * tm);
    const uint32_t mt = t % tm, nt = t / tm;
    const int m = (int)mt * 16 + (lane & 15);       // A row: output channel
    const int kcol = (int)nt * 16 + (lane & 15);     // B column
    const int st = kcol / cin, ch = kcol % cin;
    const int o[4] = {st / 27 - 1, (st / 9) % 3 - 1, (st / 3) % 3 - 1, st % 3 - 1};
    const int n[4] = {g.n1, g.n2, g.n3, g.n4};
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint32_t r0 = ch_i * chunk, r1 = min(r0 + chunk, R);
    for (uint32_t rb = r0; rb < r1; rb += 16) {
        s4 av = {0, 0, 0, 0}, bv = {0, 0, 0, 0};
        for (int q = 0; q < 4; ++q) {
            const uint32_t r = rb + 4 * (uint32_t)(lane >> 4) + (uint32_t)q;
            if (r >= r1) continue;
            av[q] = (short)gout[(size_t)r * cout + m];
            if (kcol < K) {
                const uint32_t b = r / g.V, s = r % g.V;
                int c[4];
                coords(g, s, c);
                int e[4];
                for (int d = 0; d < 4; ++d) e[d] = wrap(c[d] + o[d], n[d]);
                bv[q] = (short)in[((size_t)b * g.V + site(g, e)) * cin + ch];
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    // lane holds D[4 (lane / 16) + i][lane % 16]: rows are output channels
    for (int i = 0; i < 4; ++i) {
        const int mm = (int)mt * 16 + 4 * (lane >> 4) + i, kk = (int)nt * 16 + (lane & 15);
        if (kk < K) part[((size_t)ch_i * cout + mm) * K + kk] = acc[i];
    }
}

// out[i] = sum over the nchunk partial arrays (fixed order)
__global__ void k_flow_sum_parts(const float* part, uint32_t nchunk, uint32_t n, float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = 0.f;
    for (uint32_t c = 0; c < nchunk; ++c) a += part[(size_t)c * n + i];
    out[i] = a;
}

// bias gradient: out[c] = sum over rows of gout[r][c] (one workgroup per channel)
__global__ __launch_bounds__(256) void k_flow_colsum(const uint16_t* gout, int cout, uint32_t R, float* out)
{
    __shared__ float red[256];
    float a = 0.f;
    for (uint32_t r = threadIdx.x; r < R; r += 256) a += bf2f(gout[(size_t)r * cout + blockIdx.x]);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// transposed conv2 weights for the input gradient: wt[c][st * 16 + n] =
// W2[n][(80 - st) * 32 + c] (the stencil point of offset -o)
__global__ void k_flow_w2_transpose(const uint16_t* w2, uint16_t* wt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)(kHidden * kStencil * kOut)) return;
    const uint32_t c = i / (kStencil * kOut), rest = i % (kStencil * kOut);
    const uint32_t st = rest / kOut, nn = rest % kOut;
    wt[i] = w2[(size_t)nn * (kStencil * kHidden) + (kStencil - 1 - st) * kHidden + c];
}

// dpre = dhid (1 - h^2), h the tanh output (BF16)
__global__ void k_flow_tanh_back(const float* dhid, const uint16_t* h, uint16_t* dpre, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float hv = bf2f(h[i]);
    dpre[i] = f2bf(dhid[i] * (1.f - hv * hv));
}

// Adam (Kingma and Ba 2015) on the FP32 parameters, then the device copies
// the flow evaluates with (BF16 weights, FP32 biases) of every layer.
__global__ void k_flow_adam(float* p, const float* gr, float* m, float* v, uint32_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float gscale)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_flow_unpack_params(const float* p, uint32_t layers, uint16_t* w1, float* b1, uint16_t* w2, float* b2)
{
    const uint64_t per = (uint64_t)kHidden * kStencil * kCin0 + kHidden + (uint64_t)kOut * kStencil * kHidden + kOut;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per * layers) return;
    const uint32_t l = (uint32_t)(i / per);
    uint64_t o = i % per;
    const uint64_t n1 = (uint64_t)kHidden * kStencil * kCin0, n2 = (uint64_t)kOut * kStencil * kHidden;
    if (o < n1) { w1[l * n1 + o] = f2bf(p[i]); return; }
    o -= n1;
    if (o < kHidden) { b1[l * kHidden + o] = p[i]; return; }
    o -= kHidden;
    if (o < n2) { w2[l * n2 + o] = f2bf(p[i]); return; }
    o -= n2;
    b2[l * kOut + o] = p[i];
}

// per-site sum of (1 - (U_p)_0) over the six plaquettes with corner x
__global__ void k_flow_action_sites(const float4* L, double* out, FGeom g, uint32_t B)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B * g.V) return;
    const uint32_t b = j / g.V, s = j % g.V;
    int c[4];
    coords(g, s, c);
    double acc = 0.0;
    for (int nu = 0; nu < 4; ++nu)
        for (int rho = nu + 1; rho < 4; ++rho) acc += 1.0 - (double)plaq0(L, g, b, nu, rho, c);
    out[j] = acc;
}

// out[b] = sum of in[b * n + i] (one workgroup per row, fixed order)
template <typename T>
__global__ __launch_bounds__(256) void k_flow_rowsum(const T* in, uint32_t n, double* out)
{
    __shared__ double red[256];
    const uint32_t b = blockIdx.x;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) acc += (double)in[(size_t)b * n + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

int err(int code, const std::string& m) { return qcdhip::set_error(code, m.c_str()); }

}  // namespace

struct qcd_flow {
    qcd_flow_desc d;
    FGeom g;
    size_t nparam = 0;
    std::vector<float> host_params;
    uint16_t* w1 = nullptr;    // [layers][32][81 * 8]
    float* b1 = nullptr;       // [layers][32]
    uint16_t* w2 = nullptr;    // [layers][16][81 * 32]
    float* b2 = nullptr;       // [layers][16]
    float4* links = nullptr;   // [max_batch][4][V]
    uint16_t* feat = nullptr;  // [max_batch * V][8]
    uint16_t* hid = nullptr;   // [max_batch * V][32]
    float* prm = nullptr;      // [max_batch * V][16]
    float* logj = nullptr;     // [max_batch * V]
    double* act = nullptr;     // [max_batch * V]
    double* sums = nullptr;    // [2][max_batch]
    // learned sweeps, on the device so that a recorded graph replays them
    // with successive keys: [0] links accepted, [1] sweeps run, [2] the
    // current sweep's seed
    unsigned long long* imh = nullptr;
    // training (allocated by the first qcd_flow_train_step)
    float* tp = nullptr;        // FP32 parameters, the flat layout of qcd_flow_set_params
    float* tgrad = nullptr;     // their gradient
    float* tm = nullptr;        // Adam moments
    float* tv = nullptr;
    uint16_t* tg = nullptr;     // d loss / d spline parameters [max_batch * V][16] (BF16)
    float* tdhid = nullptr;     // d loss / d hidden [max_batch * V][32]
    uint16_t* tdpre = nullptr;  // d loss / d pre-activation (BF16)
    float* tloss = nullptr;     // -log q_a per site
    uint16_t* tw2t = nullptr;   // transposed conv2 weights of one layer
    float* tpart = nullptr;     // weight-gradient partial sums
    float* tzero = nullptr;     // zero bias of the transposed convolution
    double* tsums = nullptr;    // per layer and configuration loss
    int tstep = 0;
    bool host_stale = false;    // host_params behind tp
    hipStream_t stream = nullptr;
};

namespace {

constexpr size_t kW1 = (size_t)kHidden * kStencil * kCin0;
constexpr size_t kW2 = (size_t)kOut * kStencil * kHidden;
constexpr size_t kLayerParams = kW1 + kHidden + kW2 + kOut;

#define FCHK(x)                                                                                         \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) return err(QCD_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int run_layers(qcd_flow* F, uint32_t B, bool inverse)
{
    const FGeom g = F->g;
    const uint32_t n = B * g.V, nb = (n + 255) / 256;
    FCHK(hipMemsetAsync(F->logj, 0, sizeof(float) * n, F->stream));
    for (int step = 0; step < F->d.layers; ++step) {
        const int l = inverse ? F->d.layers - 1 - step : step;
        const int mu = l % 4, par = (l / 4) % 2;
        hipLaunchKernelGGL(k_flow_features, dim3(nb), dim3(256), 0, F->stream, F->links, F->feat, g, B, mu, par);
        const uint32_t w1n = (n / 16) * (kHidden / 16), w2n = (n / 16) * (kOut / 16);
        hipLaunchKernelGGL(k_flow_conv<kCin0>, dim3((w1n + 3) / 4), dim3(256), 0, F->stream, F->feat, kCin0, F->w1 + l * kW1,
                           F->b1 + l * kHidden, kHidden, F->hid, (float*)nullptr, g, B, 1);
        hipLaunchKernelGGL(k_flow_conv<kHidden>, dim3((w2n + 3) / 4), dim3(256), 0, F->stream, F->hid, kHidden, F->w2 + l * kW2,
                           F->b2 + l * kOut, kOut, (uint16_t*)nullptr, F->prm, g, B, 0);
        hipLaunchKernelGGL(k_flow_couple, dim3(nb), dim3(256), 0, F->stream, F->links, F->prm, F->logj, g, B, mu, par,
                           F->d.bins, inverse ? 1 : 0);
    }
    FCHK(hipGetLastError());
    return QCD_OK;
}

// per-site action terms of the configurations now in F->links
int action_sites(qcd_flow* F, uint32_t B)
{
    const uint32_t n = B * F->g.V, nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_flow_action_sites, dim3(nb), dim3(256), 0, F->stream, F->links, F->act, F->g, B);
    FCHK(hipGetLastError());
    return QCD_OK;
}

// log q from the accumulated log J, the action from the per-site terms
int finish(qcd_flow* F, uint32_t B, double beta, double* logq, double* action)
{
    const FGeom g = F->g;
    hipLaunchKernelGGL((k_flow_rowsum<float>), dim3(B), dim3(256), 0, F->stream, F->logj, g.V, F->sums);
    hipLaunchKernelGGL((k_flow_rowsum<double>), dim3(B), dim3(256), 0, F->stream, F->act, g.V, F->sums + F->d.max_batch);
    FCHK(hipGetLastError());
    std::vector<double> h(2 * (size_t)F->d.max_batch);
    FCHK(hipMemcpyAsync(h.data(), F->sums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, F->stream));
    FCHK(hipStreamSynchronize(F->stream));
    for (uint32_t b = 0; b < B; ++b) {
        if (logq) logq[b] = -h[b];
        if (action) action[b] = beta * h[F->d.max_batch + b];
    }
    return QCD_OK;
}

int check_batch(const qcd_flow* F, int batch)
{
    if (batch < 1 || batch > F->d.max_batch)
        return err(QCD_E_ARG, "batch " + std::to_string(batch) + " outside 1.." + std::to_string(F->d.max_batch));
    if (F->nparam == 0 || F->host_params.empty()) return err(QCD_E_STATE, "flow parameters not set");
    return QCD_OK;
}

}  // namespace

extern "C" {

int qcd_flow_create(const qcd_flow_desc* desc, qcd_flow** out)
{
    if (!desc || !out) return err(QCD_E_ARG, "null argument");
    const qcd_flow_desc d = *desc;
    for (int k = 0; k < 4; ++k)
        if (d.n[k] < 2) return err(QCD_E_ARG, "flow lattice extents must be >= 2");
    if (d.layers < 1 || d.layers > 64) return err(QCD_E_ARG, "layers must be 1..64");
    if (d.hidden != kHidden) return err(QCD_E_UNSUPPORTED, "hidden channels must be 32");
    if (d.bins < 1 || d.bins > kMaxBins) return err(QCD_E_ARG, "bins must be 1..5");
    if (d.max_batch < 1) return err(QCD_E_ARG, "max_batch must be >= 1");
    const uint64_t V = (uint64_t)d.n[0] * d.n[1] * d.n[2] * d.n[3];
    if (V % 16) return err(QCD_E_ARG, "the flow needs a site count divisible by 16");
    if (V * (uint64_t)d.max_batch > (1ull << 26)) return err(QCD_E_ARG, "max_batch x sites above 2^26");
    qcd_flow* F = new qcd_flow();
    F->d = d;
    F->g = FGeom{d.n[0], d.n[1], d.n[2], d.n[3], (uint32_t)V};
    F->nparam = kLayerParams * (size_t)d.layers;
    const size_t n = V * (size_t)d.max_batch;
    bool ok = hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->w1, sizeof(uint16_t) * kW1 * d.layers) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->b1, sizeof(float) * kHidden * d.layers) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->w2, sizeof(uint16_t) * kW2 * d.layers) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->b2, sizeof(float) * kOut * d.layers) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->links, sizeof(float4) * 4 * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->feat, sizeof(uint16_t) * kCin0 * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->hid, sizeof(uint16_t) * kHidden * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->prm, sizeof(float) * kOut * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->logj, sizeof(float) * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->act, sizeof(double) * n) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->sums, sizeof(double) * 2 * d.max_batch) == hipSuccess;
    ok = ok && hipMalloc((void**)&F->imh, 3 * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMemset(F->imh, 0, 3 * sizeof(unsigned long long)) == hipSuccess;
    if (!ok) {
        qcd_flow_destroy(F);
        return err(QCD_E_NOMEM, "flow buffers");
    }
    *out = F;
    return QCD_OK;
}

int qcd_flow_destroy(qcd_flow* F)
{
    if (!F) return QCD_OK;
    if (F->stream) (void)hipStreamSynchronize(F->stream);
    (void)hipFree(F->w1);
    (void)hipFree(F->b1);
    (void)hipFree(F->w2);
    (void)hipFree(F->b2);
    (void)hipFree(F->links);
    (void)hipFree(F->feat);
    (void)hipFree(F->hid);
    (void)hipFree(F->prm);
    (void)hipFree(F->logj);
    (void)hipFree(F->act);
    (void)hipFree(F->sums);
    (void)hipFree(F->imh);
    for (void* q : {(void*)F->tp, (void*)F->tgrad, (void*)F->tm, (void*)F->tv, (void*)F->tg, (void*)F->tdhid,
                    (void*)F->tdpre, (void*)F->tloss, (void*)F->tw2t, (void*)F->tpart, (void*)F->tzero,
                    (void*)F->tsums})
        (void)hipFree(q);
    if (F->stream) (void)hipStreamDestroy(F->stream);
    delete F;
    return QCD_OK;
}

int qcd_flow_param_count(const qcd_flow* F, size_t* n)
{
    if (!F || !n) return err(QCD_E_ARG, "null argument");
    *n = F->nparam;
    return QCD_OK;
}

int qcd_flow_set_params(qcd_flow* F, const float* p, size_t n)
{
    if (!F || !p) return err(QCD_E_ARG, "null argument");
    if (n != F->nparam) return err(QCD_E_ARG, "expected " + std::to_string(F->nparam) + " parameters, got " + std::to_string(n));
    const int Ln = F->d.layers;
    std::vector<uint16_t> w1(kW1 * Ln), w2(kW2 * Ln);
    std::vector<float> b1(kHidden * Ln), b2(kOut * Ln);
    for (int l = 0; l < Ln; ++l) {
        const float* q = p + l * kLayerParams;
        for (size_t i = 0; i < kW1; ++i) w1[l * kW1 + i] = f2bf(q[i]);
        for (int i = 0; i < kHidden; ++i) b1[l * kHidden + i] = q[kW1 + i];
        for (size_t i = 0; i < kW2; ++i) w2[l * kW2 + i] = f2bf(q[kW1 + kHidden + i]);
        for (int i = 0; i < kOut; ++i) b2[l * kOut + i] = q[kW1 + kHidden + kW2 + i];
    }
    FCHK(hipMemcpy(F->w1, w1.data(), sizeof(uint16_t) * w1.size(), hipMemcpyHostToDevice));
    FCHK(hipMemcpy(F->w2, w2.data(), sizeof(uint16_t) * w2.size(), hipMemcpyHostToDevice));
    FCHK(hipMemcpy(F->b1, b1.data(), sizeof(float) * b1.size(), hipMemcpyHostToDevice));
    FCHK(hipMemcpy(F->b2, b2.data(), sizeof(float) * b2.size(), hipMemcpyHostToDevice));
    F->host_params.assign(p, p + n);
    F->host_stale = false;
    if (F->tp) {   // training continues from these parameters, fresh moments
        FCHK(hipMemcpy(F->tp, p, sizeof(float) * n, hipMemcpyHostToDevice));
        FCHK(hipMemset(F->tm, 0, sizeof(float) * n));
        FCHK(hipMemset(F->tv, 0, sizeof(float) * n));
        F->tstep = 0;
    }
    return QCD_OK;
}

int qcd_flow_sample(qcd_flow* F, unsigned long long seed, int batch, double beta, float* links, double* logq,
                    double* action)
{
    if (!F || !links) return err(QCD_E_ARG, "null argument");
    int rc = check_batch(F, batch);
    if (rc) return rc;
    const uint64_t n = (uint64_t)batch * 4 * F->g.V;
    hipLaunchKernelGGL(k_flow_haar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, F->stream, F->links, n,
                       (uint64_t)seed);
    if ((rc = run_layers(F, (uint32_t)batch, false))) return rc;
    if ((rc = action_sites(F, (uint32_t)batch))) return rc;   // of the proposals
    if ((rc = finish(F, (uint32_t)batch, beta, logq, action))) return rc;
    FCHK(hipMemcpy(links, F->links, sizeof(float4) * n, hipMemcpyDeviceToHost));
    return QCD_OK;
}

int qcd_flow_evaluate(qcd_flow* F, const float* links, int batch, double beta, double* logq, double* action)
{
    if (!F || !links) return err(QCD_E_ARG, "null argument");
    int rc = check_batch(F, batch);
    if (rc) return rc;
    const uint64_t n = (uint64_t)batch * 4 * F->g.V;
    FCHK(hipMemcpyAsync(F->links, links, sizeof(float4) * n, hipMemcpyHostToDevice, F->stream));
    if ((rc = action_sites(F, (uint32_t)batch))) return rc;   // of the given configurations, before the inverse
    if ((rc = run_layers(F, (uint32_t)batch, true))) return rc;
    return finish(F, (uint32_t)batch, beta, logq, action);
}

namespace {
int imh_sweep(qcd_flow* F, qcd_lattice* lat, unsigned long long seed, int mode, double beta)
{
    if (!F || !lat) return err(QCD_E_ARG, "null argument");
    if (F->nparam == 0 || F->host_params.empty()) return err(QCD_E_STATE, "flow parameters not set");
    int n[4];
    int rc = qcdhip::lattice_dims(lat, n);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k)
        if (n[k] != F->d.n[k]) return err(QCD_E_ARG, "the lattice and the flow have different extents");
    hipStream_t st = nullptr;
    if ((rc = qcdhip::su2_quats(lat, F->links, true, &st))) return rc;   // the lattice's stream from here on
    hipLaunchKernelGGL(k_flow_imh_key, dim3(1), dim3(64), 0, st, F->imh, seed, mode);
    const FGeom g = F->g;
    const uint32_t nb = (g.V + 255) / 256;
    const uint32_t w1n = (g.V / 16) * (kHidden / 16), w2n = (g.V / 16) * (kOut / 16);
    for (int l = 0; l < F->d.layers; ++l) {
        const int mu = l % 4, par = (l / 4) % 2;
        hipLaunchKernelGGL(k_flow_features, dim3(nb), dim3(256), 0, st, F->links, F->feat, g, 1u, mu, par);
        hipLaunchKernelGGL(k_flow_conv<kCin0>, dim3((w1n + 3) / 4), dim3(256), 0, st, F->feat, kCin0, F->w1 + l * kW1,
                           F->b1 + l * kHidden, kHidden, F->hid, (float*)nullptr, g, 1u, 1);
        hipLaunchKernelGGL(k_flow_conv<kHidden>, dim3((w2n + 3) / 4), dim3(256), 0, st, F->hid, kHidden, F->w2 + l * kW2,
                           F->b2 + l * kOut, kOut, (uint16_t*)nullptr, F->prm, g, 1u, 0);
        hipLaunchKernelGGL(k_flow_imh_couple, dim3(nb), dim3(256), 0, st, F->links, F->prm, g, mu, par, F->d.bins,
                           (float)beta, ((uint64_t)l + 1) * 0xD1B54A32D192ED03ull, F->imh);
    }
    FCHK(hipGetLastError());
    return qcdhip::su2_quats(lat, F->links, false, nullptr);
}
}  // namespace

int qcd_flow_imh_sweep(qcd_flow* F, qcd_lattice* lat, unsigned long long seed, double beta)
{
    return imh_sweep(F, lat, seed, 2, beta);
}

int qcd_flow_imh_sweep_next(qcd_flow* F, qcd_lattice* lat, double beta)
{
    return imh_sweep(F, lat, 0ull, 1, beta);
}

int qcd_flow_imh_seed(qcd_flow* F, unsigned long long seed)
{
    if (!F) return err(QCD_E_ARG, "null flow");
    if (!F->imh) return err(QCD_E_STATE, "flow not set up");
    hipLaunchKernelGGL(k_flow_imh_key, dim3(1), dim3(64), 0, F->stream, F->imh, seed, 0);
    FCHK(hipStreamSynchronize(F->stream));
    return QCD_OK;
}

}  // extern "C"

namespace {

// The layer-local loss of `batch` configurations (uploaded from `links`)
// and its gradient with respect to every parameter, into F->tgrad (sums
// over the batch); *loss = the loss per configuration.
int train_grad(qcd_flow* F, const float* links, int batch, double* loss)
{
    if (!F || !links) return err(QCD_E_ARG, "null argument");
    int rc = check_batch(F, batch);
    if (rc) return rc;
    const FGeom g = F->g;
    const uint32_t B = (uint32_t)batch, R = B * g.V, nb = (R + 255) / 256;
    const size_t nmax = (size_t)F->d.max_batch * g.V;
    constexpr uint32_t kChunk = 512;
    const size_t nchunk_max = (nmax + kChunk - 1) / kChunk;
    const size_t part_max = nchunk_max * std::max((size_t)kOut * kStencil * kHidden, (size_t)kHidden * 656);
    if (!F->tp) {   // first step: training buffers and the FP32 parameters from the host copy
        bool ok = true;
        ok = ok && hipMalloc((void**)&F->tp, sizeof(float) * F->nparam) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tgrad, sizeof(float) * F->nparam) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tm, sizeof(float) * F->nparam) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tv, sizeof(float) * F->nparam) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tg, sizeof(uint16_t) * kOut * nmax) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tdhid, sizeof(float) * kHidden * nmax) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tdpre, sizeof(uint16_t) * kHidden * nmax) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tloss, sizeof(float) * nmax) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tw2t, sizeof(uint16_t) * kHidden * kStencil * kOut) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tpart, sizeof(float) * part_max) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tzero, sizeof(float) * kHidden) == hipSuccess;
        ok = ok && hipMalloc((void**)&F->tsums, sizeof(double) * F->d.layers * F->d.max_batch) == hipSuccess;
        if (!ok) return err(QCD_E_NOMEM, "flow training buffers");
        FCHK(hipMemcpy(F->tp, F->host_params.data(), sizeof(float) * F->nparam, hipMemcpyHostToDevice));
        FCHK(hipMemset(F->tm, 0, sizeof(float) * F->nparam));
        FCHK(hipMemset(F->tv, 0, sizeof(float) * F->nparam));
        FCHK(hipMemset(F->tzero, 0, sizeof(float) * kHidden));
        F->tstep = 0;
    }
    const uint64_t nl = (uint64_t)B * 4 * g.V;
    FCHK(hipMemcpyAsync(F->links, links, sizeof(float4) * nl, hipMemcpyHostToDevice, F->stream));
    const uint32_t w1n = (R / 16) * (kHidden / 16), w2n = (R / 16) * (kOut / 16);
    const uint32_t nchunk = (R + kChunk - 1) / kChunk;
    for (int l = 0; l < F->d.layers; ++l) {
        const int mu = l % 4, par = (l / 4) % 2;
        float* gl = F->tgrad + (size_t)l * kLayerParams;   // W1, b1, W2, b2 of layer l
        // forward of the conditioner on the data configurations
        hipLaunchKernelGGL(k_flow_features, dim3(nb), dim3(256), 0, F->stream, F->links, F->feat, g, B, mu, par);
        hipLaunchKernelGGL(k_flow_conv<kCin0>, dim3((w1n + 3) / 4), dim3(256), 0, F->stream, F->feat, kCin0, F->w1 + l * kW1,
                           F->b1 + l * kHidden, kHidden, F->hid, (float*)nullptr, g, B, 1);
        hipLaunchKernelGGL(k_flow_conv<kHidden>, dim3((w2n + 3) / 4), dim3(256), 0, F->stream, F->hid, kHidden, F->w2 + l * kW2,
                           F->b2 + l * kOut, kOut, (uint16_t*)nullptr, F->prm, g, B, 0);
        // -log q_a and its gradient with respect to the spline parameters
        hipLaunchKernelGGL(k_flow_nll_grad, dim3(nb), dim3(256), 0, F->stream, (const float4*)F->links, F->prm, F->tg,
                           F->tloss, g, B, mu, par, F->d.bins);
        hipLaunchKernelGGL((k_flow_rowsum<float>), dim3(B), dim3(256), 0, F->stream, F->tloss, g.V,
                           F->tsums + (size_t)l * F->d.max_batch);
        // conv2: weight and bias gradients, then the hidden gradient through the transposed convolution
        {
            const uint32_t tiles = (uint32_t)(kStencil * kHidden / 16) * (kOut / 16) * nchunk;
            hipLaunchKernelGGL(k_flow_wgrad, dim3((tiles + 3) / 4), dim3(256), 0, F->stream, (const uint16_t*)F->tg, kOut,
                               (const uint16_t*)F->hid, kHidden, F->tpart, g, B, kChunk);
            const uint32_t n2 = (uint32_t)kW2;
            hipLaunchKernelGGL(k_flow_sum_parts, dim3((n2 + 255) / 256), dim3(256), 0, F->stream, (const float*)F->tpart,
                               nchunk, n2, gl + kW1 + kHidden);
            hipLaunchKernelGGL(k_flow_colsum, dim3(kOut), dim3(256), 0, F->stream, (const uint16_t*)F->tg, kOut, R,
                               gl + kW1 + kHidden + kW2);
            const uint32_t nt = (uint32_t)(kHidden * kStencil * kOut);
            hipLaunchKernelGGL(k_flow_w2_transpose, dim3((nt + 255) / 256), dim3(256), 0, F->stream,
                               (const uint16_t*)(F->w2 + l * kW2), F->tw2t);
            const uint32_t wtn = (R / 16) * (kHidden / 16);
            hipLaunchKernelGGL(k_flow_conv<kOut>, dim3((wtn + 3) / 4), dim3(256), 0, F->stream, (const uint16_t*)F->tg, kOut,
                               (const uint16_t*)F->tw2t, (const float*)F->tzero, kHidden, (uint16_t*)nullptr, F->tdhid,
                               g, B, 0);
            hipLaunchKernelGGL(k_flow_tanh_back, dim3((R * kHidden + 255) / 256), dim3(256), 0, F->stream,
                               (const float*)F->tdhid, (const uint16_t*)F->hid, F->tdpre, R * (uint32_t)kHidden);
        }
        // conv1: weight and bias gradients (the features are data: no further gradient)
        {
            const uint32_t K1 = kStencil * kCin0, tiles = ((K1 + 15) / 16) * (kHidden / 16) * nchunk;
            hipLaunchKernelGGL(k_flow_wgrad, dim3((tiles + 3) / 4), dim3(256), 0, F->stream, (const uint16_t*)F->tdpre,
                               kHidden, (const uint16_t*)F->feat, kCin0, F->tpart, g, B, kChunk);
            const uint32_t n1 = (uint32_t)kW1;
            hipLaunchKernelGGL(k_flow_sum_parts, dim3((n1 + 255) / 256), dim3(256), 0, F->stream, (const float*)F->tpart,
                               nchunk, n1, gl);
            hipLaunchKernelGGL(k_flow_colsum, dim3(kHidden), dim3(256), 0, F->stream, (const uint16_t*)F->tdpre, kHidden,
                               R, gl + kW1);
        }
    }
    FCHK(hipGetLastError());
    if (loss) {
        std::vector<double> h((size_t)F->d.layers * F->d.max_batch);
        FCHK(hipMemcpyAsync(h.data(), F->tsums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, F->stream));
        FCHK(hipStreamSynchronize(F->stream));
        double a = 0.0;
        for (int l = 0; l < F->d.layers; ++l)
            for (uint32_t b = 0; b < B; ++b) a += h[(size_t)l * F->d.max_batch + b];
        *loss = a / B;
    }
    return QCD_OK;
}

}  // namespace

extern "C" {

int qcd_flow_train_grad(qcd_flow* F, const float* links, int batch, double* loss, float* grad)
{
    const int rc = train_grad(F, links, batch, loss);
    if (rc) return rc;
    if (grad) FCHK(hipMemcpy(grad, F->tgrad, sizeof(float) * F->nparam, hipMemcpyDeviceToHost));
    return QCD_OK;
}

int qcd_flow_train_step(qcd_flow* F, const float* links, int batch, float lr, double* loss)
{
    if (!(lr > 0.f)) return err(QCD_E_ARG, "learning rate must be positive");
    const int rc = train_grad(F, links, batch, loss);
    if (rc) return rc;
    const uint32_t B = (uint32_t)batch;
    // Adam on the mean over the batch, then the flow's own copies
    ++F->tstep;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)F->tstep), c2 = 1.f - powf(b2, (float)F->tstep);
    const uint32_t np = (uint32_t)F->nparam;
    hipLaunchKernelGGL(k_flow_adam, dim3((np + 255) / 256), dim3(256), 0, F->stream, F->tp, (const float*)F->tgrad,
                       F->tm, F->tv, np, lr, b1, b2, 1e-8f, c1, c2, 1.f / (float)B);
    hipLaunchKernelGGL(k_flow_unpack_params, dim3((np + 255) / 256), dim3(256), 0, F->stream, (const float*)F->tp,
                       (uint32_t)F->d.layers, F->w1, F->b1, F->w2, F->b2);
    FCHK(hipGetLastError());
    F->host_stale = true;
    return QCD_OK;
}

int qcd_flow_get_params(qcd_flow* F, float* p, size_t n)
{
    if (!F || !p) return err(QCD_E_ARG, "null argument");
    if (n != F->nparam) return err(QCD_E_ARG, "expected " + std::to_string(F->nparam) + " parameters, got " + std::to_string(n));
    if (F->host_stale) {
        FCHK(hipStreamSynchronize(F->stream));
        FCHK(hipMemcpy(F->host_params.data(), F->tp, sizeof(float) * F->nparam, hipMemcpyDeviceToHost));
        F->host_stale = false;
    }
    if (F->host_params.size() != n) return err(QCD_E_STATE, "flow parameters not set");
    memcpy(p, F->host_params.data(), sizeof(float) * n);
    return QCD_OK;
}

int qcd_flow_imh_counts(qcd_flow* F, unsigned long long* accepted, unsigned long long* proposed, int reset)
{
    if (!F) return err(QCD_E_ARG, "null flow");
    FCHK(hipDeviceSynchronize());
    unsigned long long a[2] = {0, 0};
    FCHK(hipMemcpy(a, F->imh, sizeof(a), hipMemcpyDeviceToHost));
    if (accepted) *accepted = a[0];
    if (proposed) *proposed = a[1] * (unsigned long long)F->d.layers * (F->g.V / 2);   // every layer proposes V/2 links
    if (reset) FCHK(hipMemset(F->imh, 0, 2 * sizeof(unsigned long long)));   // the seed stays
    return QCD_OK;
}

}  // extern "C"
