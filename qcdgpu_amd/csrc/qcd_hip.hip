// qcd_hip.hip -- gfx950 kernels and the C ABI (include/qcd_hip.h) for the
// SU(N) heatbath / overrelaxation update and the plaquette / Polyakov
// measurements of QCDGPU.
//
// Reference path replaced (paths relative to the reference's SU(N)/ tree):
//   update_{even,odd}_{X,Y,Z,T}   suncl/suncl.cl:399-975
//     lattice_staple_3/_2         suncl/su3_update_cl.cl:341-565, su2_update_cl.cl:206-431
//     lattice_heatbath2/3         suncl/su3_update_cl.cl:568-779, su2_update_cl.cl:434-550
//   PRNG produce (xor128, pm)     random/prngcl_xor128.cl:100, prngcl_pm.cl:86
//   lattice_init_cold / _gid      suncl/suncl.cl:164-347
//   lattice_GramSchmidt           suncl/suncl.cl:350-396
//   lattice_measurement_plq       suncl/sun_measurements_cl.cl:1178-1323
//   lattice_polyakov (+ reduce)   suncl/polyakov.cl:54-141, 405-431
//
// What is different by design:
//   * one thread per link of the (parity, mu) half-step reads a
//     checkerboard-compacted plane-major table (qcd_device.hpp), so every
//     neighbour load of a wavefront is a contiguous run;
//   * the PRN buffer is gone: each thread owns one generator stream and draws
//     in registers; in reference stream mode it then skips to the next
//     multiple of S quads, which reproduces the buffer the reference kernel
//     would have read (instance i = thread i, PRNG_instances = PRNGSTEP);
//   * reunitarisation is fused into every link update and done in FP64;
//   * measurements reduce in FP64 with a deterministic two-level tree (wave
//     shuffles, LDS, then the last workgroup sums the per-workgroup partials in
//     index order), appending one record per call on the device;
//   * a whole measurement cycle is captured once into a hipGraph.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/qcd_hip.h"
#include "qcd_device.hpp"
#include "qcd_internal.h"

using namespace qcd;

namespace {

constexpr int kBlock = 256;
constexpr size_t kScratchBytesDev = (size_t)16 << 20;   // the allocation-free measurement's device array
constexpr double kPi2 = 6.2831853071795864769252867665590;

// RK_BUF: quads read from a buffer filled by a produce kernel (RANLUX),
// quad k of stream i at buf[k * n + i] as the reference's PRN buffer.
enum RngKind { RK_NONE = 0, RK_GID = 1, RK_XOR128 = 2, RK_PM = 3, RK_BUF = 4 };

struct UpdArgs {
    void* table;
    // Table the staple gathers read: the table itself, or (shadow mode, a
    // race check) a copy of it taken before the half-step.  A half-step
    // reads only links it does not write, so both give the same bits.
    const void* src;
    Geom g;
    int parity;
    int mu;
    double beta;          // BETA / N
    int nhit;
    uint32_t sites;       // V, for GID mode
    uint4* rng;           // per-stream state at the start of the half-step (stream = compact index)
    uint4* rng_out;       // state after the half-step (the two buffers alternate)
    const uint4* jump;    // XOR128: jump matrix for 4 * s_quads steps (128 columns)
    const float4* prn;    // RK_BUF: quad k of stream i at prn[k * prn_stride + i]
    uint32_t prn_stride;
    uint32_t pm_ak;       // Park-Miller: 16807^s_quads mod (2^31 - 1)
    uint32_t s_quads;     // quads per heatbath half-step per stream (0 = fast)
    // fused plaquette measurement (last sweep of a measurement cycle):
    // partials[(meas_mode - 1) * gridDim.x + block] = sum over the block of
    // ReTr(U' * sum_{nu<mu} S_nu); modes 1, 2, 3 = even Y, Z, T half-steps.
    double* partials;
    int meas_mode;
    int xcd_map;          // 0: hardware order, 1: one x-slab per XCD, >= 2: blocks of (x, z) per XCD
    // xcd_map >= 2: z-groups of the lattice, N4, x-values per XCD block,
    // XCD blocks along x, z-groups per XCD block
    uint32_t ord_ng, ord_nt, ord_xs, ord_ax, ord_zr;
    uint32_t unit0;       // first 64-link unit of the launch (an x-slab's first interior slice; else 0)
    Twist tw;             // twisted boundary (read only by the TW instantiations)
};

// 64-link unit handled by the l-th workgroup (or wavefront) dispatched to
// XCD k, for xcd_map >= 2.  A unit is a group of rb = 64 / (N2 / 2) whole
// z-rows, so unit u = zg + ng (t + N4 x).  XCD k = kx + ax kz owns the block
// x in [kx xs, (kx+1) xs), z-groups [kz zr, (kz+1) zr); within it the units
// go x offset fastest, then t, then z-group, so a unit's x, t and z
// neighbours are dispatched close to it.
__device__ __forceinline__ uint32_t block_unit(const UpdArgs& a, uint32_t k, uint32_t l)
{
    const uint32_t kx = k % a.ord_ax, kz = k / a.ord_ax;
    const uint32_t c = l % a.ord_xs, r = l / a.ord_xs;
    const uint32_t t = r % a.ord_nt, zl = r / a.ord_nt;
    return kz * a.ord_zr + zl + a.ord_ng * (t + a.ord_nt * (kx * a.ord_xs + c));
}

// Hardware dispatches workgroup b to XCD (b mod 8).  Remap so that each XCD
// works on one contiguous range of compact indices (an x-slab of the
// lattice): z/t/x neighbours then mostly hit the same XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    if (nb % 8u) return b;
    return (b % 8u) * (nb / 8u) + (b / 8u);
}

// ---------------------------------------------------------------- RNG sources
struct SrcGid {
    double gid, fs;
    int nhit;
    __device__ float4 try_quad()
    {
        // suncpp/su2/update_su2.cpp:71-76: the four numbers of a try, rounded
        // to float as suncpp does (the OpenCL FP64 build keeps them double).
        double c = 0.005 * (1.0 + nhit);
        return make_float4(fabsf((float)sin((c + 270.0 / fs) * gid)), fabsf((float)cos((c + 60.0 / fs) * gid)),
                           fabsf((float)sin((c - 150.0 / fs) * gid)), fabsf((float)cos((c - 380.0 / fs) * gid)));
    }
    __device__ float4 accept_quad()
    {
        return make_float4(fabsf((float)cos((0.08 - 270.0 / fs) * gid)), fabsf((float)sin((0.08 - 60.0 / fs) * gid)),
                           0.f, 0.f);
    }
    __device__ void skip() {}
};

struct SrcXor {
    uint4 s;
    uint32_t used;
    __device__ float4 try_quad()
    {
        used++;
        return xor128_quad(s);
    }
    __device__ float4 accept_quad() { return try_quad(); }
    __device__ void skip()
    {
        used++;
        (void)xor128_quad(s);
    }
};

struct SrcPm {
    uint4 s;
    uint32_t used;
    __device__ float4 try_quad()
    {
        used++;
        return pm_quad(s);
    }
    __device__ float4 accept_quad() { return try_quad(); }
    __device__ void skip()
    {
        used++;
        (void)pm_quad(s);
    }
};
struct SrcBuf {
    const float4* buf;
    uint32_t n, i, k;
    __device__ float4 try_quad() { return buf[(size_t)(k++) * n + i]; }
    __device__ float4 accept_quad() { return try_quad(); }
    __device__ void skip() { ++k; }
};

// ---------------------------------------------------------------- KP / OR
// b: 2x2 complex block (b00, b01, b10, b11) as 8 values of the compute type
// T (float for FP32 links, double for FP64 links).  out = c * aH with
// aH = (quaternion part of b / its norm)^dagger and c drawn from the
// Kennedy-Pendleton density (su3_update_cl.cl:568-669 with the argument
// order of suncpp/su2/update_su2.cpp:40-126), or c = aH for the
// overrelaxation reflection.  Returns false if all NHIT tries were rejected.
template <typename T> __device__ __forceinline__ void sincos_t(T x, T* s, T* c)
{
    if constexpr (sizeof(T) == 4) sincosf(x, s, c);
    else sincos(x, s, c);
}

template <typename T, bool OVER, int RK, typename Src>
__device__ __forceinline__ bool kp_update(const T* b, T beta, int nhit, Src& src, T* out)
{
    const T Mx = b[0] + b[6];
    const T My = b[1] - b[7];
    const T Mz = b[2] - b[4];
    const T Mw = b[3] + b[5];
    const T n2 = Mx * Mx + My * My + Mz * Mz + Mw * Mw;
    T id;
    if constexpr (sizeof(T) == 4) id = __builtin_amdgcn_rsqf(n2);   // one ulp
    else id = rsqrt_d(n2);
    const T det = n2 * id;
    const T a00r = Mx * id, a00i = -My * id, a01r = -Mz * id, a01i = -Mw * id;
    // aH = [[a00, a01], [-conj(a01), conj(a00)]]
    T c00r, c00i, c01r, c01i;
    if constexpr (OVER) {
        c00r = a00r;
        c00i = a00i;
        c01r = a01r;
        c01i = a01i;
    } else {
        const T pi2 = (T)kPi2;
        const T bdet = beta * det;
        const T ibdet = T(1) / bdet;
        const T ibdet_ln2 = ibdet * (T)0.69314718055994530942;
        bool flag = false;
        T delta = T(0);
        for (int i = 0; i < nhit && !flag; ++i) {
            float4 r = src.try_quad();
            T cosrnd;
            if constexpr (sizeof(T) == 4) {
                float sn;
                sincos2pi_f(r.y, &sn, &cosrnd);
            } else {
                cosrnd = cos(pi2 * (T)r.y);
            }
            if constexpr (RK == RK_GID) {
                delta = cosrnd * cosrnd * ibdet;
                flag = true;
            } else {
                if constexpr (sizeof(T) == 4) {
                    // ln(u) = log2(u) ln 2 with the hardware log2 (one ulp; 1 - r
                    // is a normal float or 0, which gives -inf and a rejection
                    // exactly as ln(0) does).
                    const float l1 = __builtin_amdgcn_logf(1.0f - r.x);
                    const float l2 = __builtin_amdgcn_logf(1.0f - r.z);
                    delta = -fmaf(cosrnd * cosrnd, l2, l1) * ibdet_ln2;
                } else {
                    delta = -(log(T(1) - (T)r.x) + cosrnd * cosrnd * log(T(1) - (T)r.z)) * ibdet;
                }
                float w2 = r.w * r.w;
                flag = ((T)w2 <= (T(1) - T(0.5) * delta));
            }
        }
        if (!flag) {
            src.skip();   // su3_update_cl.cl:665-668: an all-rejected update skips one quad
            return false;
        }
        float4 r = src.accept_quad();
        // a0 = cos(alpha) = 1 - delta, cos(theta) = 2 r.x - 1, phi = 2 pi r.y.
        // The sines are formed without the cancellation of sqrt(1 - c^2)
        // near |c| = 1: sin(alpha) = sqrt(delta (2 - delta)) and
        // sin(theta) = 2 sqrt(r.x (1 - r.x)), so FP32 keeps full relative
        // precision; FP64 links form them in double as sqrt(1 - c^2).
        if constexpr (sizeof(T) == 4) {
            const float costh = fmaf(2.0f, r.x, -1.0f);
            const float sinth = 2.0f * __builtin_amdgcn_sqrtf(r.x * (1.0f - r.x));
            const float sinal = __builtin_amdgcn_sqrtf(delta * (2.0f - delta));
            float sinphi, cosphi;
            sincos2pi_f(r.y, &sinphi, &cosphi);
            c00r = 1.0f - delta;
            c00i = sinal * costh;
            c01r = (sinal * sinth) * sinphi;
            c01i = (sinal * sinth) * cosphi;
        } else {
            const double cosal = 1.0 - (double)delta;
            const double costh = 2.0 * (double)r.x - 1.0;
            const double sinth = sqrt(1.0 - costh * costh);
            const double sinal = sqrt(1.0 - cosal * cosal);
            T sinphi, cosphi;
            sincos_t(pi2 * (T)r.y, &sinphi, &cosphi);
            c00r = (T)cosal;
            c00i = (T)(sinal * costh);
            c01r = (T)(sinal * sinth) * sinphi;
            c01i = (T)(sinal * sinth) * cosphi;
        }
    }
    // out = c * aH with c = [[c00, c01], [-conj(c01), conj(c00)]]
    const T a10r = -a01r, a10i = a01i, a11r = a00r, a11i = -a00i;
    const T c10r = -c01r, c10i = c01i, c11r = c00r, c11i = -c00i;
    out[0] = c00r * a00r - c00i * a00i + c01r * a10r - c01i * a10i;
    out[1] = c00r * a00i + c00i * a00r + c01r * a10i + c01i * a10r;
    out[2] = c00r * a01r - c00i * a01i + c01r * a11r - c01i * a11i;
    out[3] = c00r * a01i + c00i * a01r + c01r * a11i + c01i * a11r;
    out[4] = c10r * a00r - c10i * a00i + c11r * a10r - c11i * a10i;
    out[5] = c10r * a00i + c10i * a00r + c11r * a10i + c11i * a10r;
    out[6] = c10r * a01r - c10i * a01i + c11r * a11r - c11i * a11i;
    out[7] = c10r * a01i + c10i * a01r + c11r * a11i + c11i * a11r;
    return true;
}

// Left-multiply rows a, b of X by the 2x2 matrix r.
template <typename R>
__device__ __forceinline__ void apply_rows(M<R, 3>& X, int a, int b, const R* r)
{
    const C<R> r00 = {r[0], r[1]}, r01 = {r[2], r[3]};
    const C<R> r10 = {r[4], r[5]}, r11 = {r[6], r[7]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const C<R> xa = X.e[a][k], xb = X.e[b][k];
        C<R> na = cmul(r00, xa), nb = cmul(r10, xa);
        cfma(na, r01, xb);
        cfma(nb, r11, xb);
        X.e[a][k] = na;
        X.e[b][k] = nb;
    }
}

// Entry (r, c) of U * A.
template <typename R>
__device__ __forceinline__ C<R> ua_entry(const M<R, 3>& U, const M<R, 3>& A, int r, int c)
{
    C<R> s = cmul(U.e[r][0], A.e[0][c]);
    cfma(s, U.e[r][1], A.e[1][c]);
    cfma(s, U.e[r][2], A.e[2][c]);
    return s;
}

// ---------------------------------------------------------------- staple
// One direction's share of the staple sum,
//   S_nu = U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
// (suncpp/Update/sun_update.h:44-77; su3_update_cl.cl:341-565).  nu, mu and
// the parity p are wavefront-uniform.  TW: apply the twisted boundary (a
// separate instantiation, so the untwisted kernels carry none of its code).
template <typename R, int N, int NREAL, bool TW = false>
__device__ __forceinline__ M<R, N> staple_nu(const Table<R, N, NREAL>& T, const Geom& g, Site s, uint32_t i, int p,
                                             int mu, int nu, const Twist& tw = Twist{})
{
    using Tab = Table<R, N, NREAL>;
    const int q = p ^ 1;
    // The six gathers go out as two batches of three (forward, backward
    // term): three matrices in flight per wavefront keep the register count
    // low enough for six to eight wavefronts per SIMD, which hides more
    // latency than a single batch of six at four wavefronts per SIMD.
    const int cmu = coord(s, mu), cnu = coord(s, nu);
    const uint32_t o = (uint32_t)s.y & 1u;
    const uint32_t i_pmu = step_fwd(g, i, mu, cmu, o);
    const uint32_t i_pnu = step_fwd(g, i, nu, cnu, o);
    const uint32_t i_mnu = step_bwd(g, i, nu, cnu, o);
    const uint32_t i_pmu_mnu = step_bwd(g, i_pmu, nu, cnu, o);
    M<R, N> S;
    if constexpr (N == 3 && NREAL == 18 && sizeof(R) == 4) {
        // FP32, 18 reals: the forward and then the backward three gathers as
        // two batches, scheduled freely (the peak holds three matrices in
        // flight, 78 VGPRs: six wavefronts per SIMD instead of four).
        M<R, N> f2 = T.load_raw_u(q, mu, i_pnu);
        M<R, N> f1 = T.load_raw_u(q, nu, i_pmu);
        M<R, N> f3 = T.load_raw_u(p, nu, i);
        Tab::finish(f1);
        Tab::finish(f2);
        Tab::finish(f3);
        M<R, N> b1 = T.load_raw_u(q, mu, i_mnu);
        M<R, N> b3 = T.load_raw_u(q, nu, i_mnu);
        M<R, N> b2 = T.load_raw_u(p, nu, i_pmu_mnu);
        Tab::finish(b1);
        Tab::finish(b2);
        Tab::finish(b3);
        if constexpr (TW) {
            if (mu == 1 || nu == 1) {   // wavefront-uniform
                tbc(f2, tw, mu, i_pnu);
                tbc(b1, tw, mu, i_mnu);
                tbc(f1, tw, nu, i_pmu);
                tbc(b3, tw, nu, i_mnu);
                tbc(f3, tw, nu, i);
                tbc(b2, tw, nu, i_pmu_mnu);
            }
        }
        const M<R, 3> F = su3_fwd_term(f1, f2, f3);
        const M<R, 3> B = su3_bwd_term(b1, b2, b3);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) S.e[r][c] = cadd(F.e[r][c], B.e[r][c]);
    } else {
        // Other storages: the forward term is gathered and formed before the
        // backward gathers are issued (the scheduling barrier keeps the two
        // apart), so at most three matrices and one partial sum are live.
        {
            M<R, N> f2 = T.load_raw_u(q, mu, i_pnu);
            M<R, N> f1 = T.load_raw_u(q, nu, i_pmu);
            M<R, N> f3 = T.load_raw_u(p, nu, i);
            Tab::finish(f1);
            Tab::finish(f2);
            Tab::finish(f3);
            if constexpr (TW) {
                if (mu == 1 || nu == 1) {
                    tbc(f2, tw, mu, i_pnu);
                    tbc(f1, tw, nu, i_pmu);
                    tbc(f3, tw, nu, i);
                }
            }
            if constexpr (N == 3) {
                S = su3_fwd_term(f1, f2, f3);
            } else {
                m_zero(S);
                const M<R, N> tf = m_mul_nd(f1, f2);
                m_acc_nd(S, tf, f3);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            M<R, N> b1 = T.load_raw_u(q, mu, i_mnu);
            M<R, N> b3 = T.load_raw_u(q, nu, i_mnu);
            M<R, N> b2 = T.load_raw_u(p, nu, i_pmu_mnu);
            Tab::finish(b1);
            Tab::finish(b2);
            Tab::finish(b3);
            if constexpr (TW) {
                if (mu == 1 || nu == 1) {
                    tbc(b1, tw, mu, i_mnu);
                    tbc(b3, tw, nu, i_mnu);
                    tbc(b2, tw, nu, i_pmu_mnu);
                }
            }
            if constexpr (N == 3) {
                const M<R, 3> B = su3_bwd_term(b1, b2, b3);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) S.e[r][c] = cadd(S.e[r][c], B.e[r][c]);
            } else {
                const M<R, N> tb = m_mul(b1, b2);
                m_acc_dn(S, tb, b3);
            }
        }
    }
    return S;
}

// ---------------------------------------------------------------- fused measurement
// Reductions avoid device-scope atomics (on MI355X they serialise across the
// eight XCDs): every workgroup writes its partial sum to its own slot and one
// single-workgroup kernel (k_finalize) sums the slots in index order.
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// ---------------------------------------------------------------- update kernel
// Workgroup = 3 wavefronts for 64 links of one (parity, mu) half-step.
// Wavefront w computes the staple share of the w-th direction nu != mu
// (uniform per wavefront, so no divergence); shares 1 and 2 pass through LDS
// to wavefront 0, which sums them, updates the link (heatbath or
// overrelaxation in the three SU(2) subgroups, FP64 Gram-Schmidt), stores it
// and, for the last even Y/Z/T half-steps of a measurement cycle, adds
// ReTr(U' * sum_{nu<mu} S_nu) to the plaquette record.
constexpr int kLinks = 64;
constexpr int kUpdThreads = 3 * kLinks;

// The link update of one active lane: U (the link as loaded), the three
// SU(2) subgroups (heatbath or overrelaxation; SU(2): one), FP64
// reunitarisation, store.  Returns ReTr(U' * Alow) for the fused plaquette
// (0 unless MEAS).  `rs` is the generator state at the start of the
// half-step; in the fast stream layout the state after consumption is
// written to rng_out.
template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS, bool TW = false>
__device__ __forceinline__ double update_link(const UpdArgs& a, const Table<R, N, NREAL>& Tb, uint32_t i, const Site& s,
                                              const M<R, N>& A, const M<R, N>& Alow, uint4 rs, M<R, N>& U)
{
    using T = R;
    const int p = a.parity, mu = a.mu;
    double trace = 0.0;
    if constexpr (TW) tbc(U, a.tw, mu, i);   // stored back as updated, the twist kept (as the reference)
    SrcGid sg{};
    SrcXor sx{};
    SrcPm sp{};
    SrcBuf sb{};
    if constexpr (RK == RK_BUF) sb = SrcBuf{a.prn, a.prn_stride, i, 0u};
    if constexpr (RK == RK_GID) {
        sg.gid = (double)((uint32_t)s.y + (uint32_t)a.g.n2 * ((uint32_t)s.z + (uint32_t)a.g.n3 *
                                                                 ((uint32_t)s.t + (uint32_t)a.g.n4 * (uint32_t)s.x)));
        sg.fs = (double)a.sites;
        sg.nhit = a.nhit;
    } else if constexpr (RK == RK_XOR128) {
        sx.s = rs;
        sx.used = 0;
    } else if constexpr (RK == RK_PM) {
        sp.s = rs;
        sp.used = 0;
    }
    const T beta = (T)a.beta;

    T blk8[8], r[8];
    if constexpr (N == 3) {
#pragma unroll
        for (int sgi = 0; sgi < 3; ++sgi) {
            const int ra = (sgi == 2) ? 1 : 0;
            const int rb = (sgi == 0) ? 1 : 2;
            const C<R> waa = ua_entry(U, A, ra, ra), wab = ua_entry(U, A, ra, rb);
            const C<R> wba = ua_entry(U, A, rb, ra), wbb = ua_entry(U, A, rb, rb);
            blk8[0] = waa.re; blk8[1] = waa.im; blk8[2] = wab.re; blk8[3] = wab.im;
            blk8[4] = wba.re; blk8[5] = wba.im; blk8[6] = wbb.re; blk8[7] = wbb.im;
            bool ok;
            if constexpr (RK == RK_GID) ok = kp_update<T, OVER, RK>(blk8, beta, a.nhit, sg, r);
            else if constexpr (RK == RK_XOR128) ok = kp_update<T, OVER, RK>(blk8, beta, a.nhit, sx, r);
            else if constexpr (RK == RK_PM) ok = kp_update<T, OVER, RK>(blk8, beta, a.nhit, sp, r);
            else if constexpr (RK == RK_BUF) ok = kp_update<T, OVER, RK>(blk8, beta, a.nhit, sb, r);
            else ok = kp_update<T, OVER, RK>(blk8, beta, a.nhit, sg, r);
            if (!ok) {
                // lattice_unity_2 (su3_update_cl.cl:665): identity in this subgroup
                r[0] = T(1); r[1] = T(0); r[2] = T(0); r[3] = T(0);
                r[4] = T(0); r[5] = T(0); r[6] = T(1); r[7] = T(0);
            }
            apply_rows(U, ra, rb, r);
        }
        M<double, 3> Ud = m_cast<double>(U);
        su3_gram_schmidt(Ud);
        U = m_cast<R>(Ud);
    } else {
        if constexpr (OVER) {
            const M<R, 2> W = m_mul(U, A);
            blk8[0] = W.e[0][0].re; blk8[1] = W.e[0][0].im; blk8[2] = W.e[0][1].re; blk8[3] = W.e[0][1].im;
            blk8[4] = W.e[1][0].re; blk8[5] = W.e[1][0].im; blk8[6] = W.e[1][1].re; blk8[7] = W.e[1][1].im;
            kp_update<T, true, RK>(blk8, beta, a.nhit, sg, r);
            M<R, 2> Rm;
            Rm.e[0][0] = {r[0], r[1]};
            Rm.e[0][1] = {r[2], r[3]};
            Rm.e[1][0] = {r[4], r[5]};
            Rm.e[1][1] = {r[6], r[7]};
            U = m_mul(Rm, U);
        } else {
            blk8[0] = A.e[0][0].re; blk8[1] = A.e[0][0].im; blk8[2] = A.e[0][1].re; blk8[3] = A.e[0][1].im;
            blk8[4] = A.e[1][0].re; blk8[5] = A.e[1][0].im; blk8[6] = A.e[1][1].re; blk8[7] = A.e[1][1].im;
            bool ok;
            if constexpr (RK == RK_GID) ok = kp_update<T, false, RK>(blk8, beta, a.nhit, sg, r);
            else if constexpr (RK == RK_XOR128) ok = kp_update<T, false, RK>(blk8, beta, a.nhit, sx, r);
            else if constexpr (RK == RK_BUF) ok = kp_update<T, false, RK>(blk8, beta, a.nhit, sb, r);
            else ok = kp_update<T, false, RK>(blk8, beta, a.nhit, sp, r);
            if (ok) {   // a rejected SU(2) update keeps the old link (su2_update_cl.cl:542-547)
                U.e[0][0] = {r[0], r[1]};
                U.e[0][1] = {r[2], r[3]};
                U.e[1][0] = {r[4], r[5]};
                U.e[1][1] = {r[6], r[7]};
            }
        }
        M<double, 2> Ud = m_cast<double>(U);
        su2_normalize(Ud);
        U = m_cast<R>(Ud);
    }
    Tb.store_u(p, mu, i, U);
    if constexpr (MEAS) trace = m_retr_mul(U, Alow);

    if (a.s_quads == 0) {
        if constexpr (RK == RK_XOR128) a.rng_out[i] = sx.s;
        else if constexpr (RK == RK_PM) a.rng_out[i] = sp.s;
    }
    return trace;
}

// Workgroup -> 64-link block.  Mode 1: XCD k (hardware block b goes to XCD
// b mod 8) gets the k-th contiguous eighth of the compact index range, an
// x-slab of the lattice.  Modes 2 and 3: each XCD gets one block of x-values
// and z-groups (block_unit; mode 2 the x-slabs of mode 1, mode 3 the block
// shape with the fewest halo links), dispatched cross-section first.
__device__ __forceinline__ uint32_t upd_block(const UpdArgs& a)
{
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    if (a.xcd_map == 0 || (nb % 8u)) return b;
    const uint32_t per = nb / 8u, k = b % 8u, l = b / 8u;
    if (a.xcd_map >= 2) return block_unit(a, k, l);
    return k * per + l;
}

// Body of the three-wavefront update for 64-link unit `blk` of a lattice
// with `nunits` units (the fused-plaquette partial of the unit goes to
// partials[(meas_mode - 1) * nunits + blk]).
// MU >= 0: the direction is a compile-time constant (the neighbour steps
// and plane bases fold away, shortening the scalar prologue before the
// gathers); MU = -1 reads it from the arguments.
template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS, bool TW = false, int MU = -1>
__device__ __forceinline__ void update3(const UpdArgs& a, uint32_t blk, uint32_t nunits)
{
    using T = R;   // compute type of the SU(2) subgroup arithmetic
    constexpr int NC = 2 * N * N;   // reals per matrix
    // shares of wavefronts 1 and 2; with MEAS also wavefront 0's own, so
    // that no staple share stays in registers through the link update
    __shared__ R lds[MEAS ? 3 : 2][NC][kLinks];
    // Wavefront index, made visibly uniform so that nu and every plane base
    // derived from it live in scalar registers.
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const uint32_t i = (a.unit0 + blk) * kLinks + lane;
    const bool active = i < a.g.vh;
    Table<R, N, NREAL> Tb{(typename Vec2<R>::T*)a.table, a.g.vh_pad};
    const Table<R, N, NREAL> Ts{(typename Vec2<R>::T*)a.src, a.g.vh_pad};
    const int p = a.parity, mu = (MU >= 0) ? MU : a.mu;
    const int nu = w + (w >= mu ? 1 : 0);
    const Site s = site_of(a.g, active ? i : 0u, p);

    // Generator state at the start of the half-step: wavefront 0 consumes
    // it.  In the reference stream layout every stream then advances by
    // exactly s_quads quads (the reference's produce() of the whole buffer):
    // after the barrier, while wavefront 0 updates the links, wavefronts 1
    // and 2 jump the start state ahead, each computing two of its four
    // words, into the other state buffer.  In the fast layout wavefront 0
    // writes the state it ends with.
    constexpr bool kStream = (RK == RK_XOR128 || RK == RK_PM);
    uint4 rs = make_uint4(0, 0, 0, 0);
    if constexpr (kStream) {
        if ((w == 0 && active) || (w > 0 && a.s_quads)) rs = a.rng[i];
    }
    // XOR128 jump table (8 KiB) staged in LDS by the whole workgroup with
    // direct global -> LDS loads (global_load_lds_dwordx4: lane l of a
    // wavefront writes 16 bytes at the wave-uniform LDS base + 16 l), so no
    // register holds it and no wait for it comes before the staple gathers
    // are issued; the vmcnt wait and the barrier below publish it.
    __shared__ uint4 jtab[RK == RK_XOR128 ? kJumpTab : 1];
    if constexpr (RK == RK_XOR128) {
        static_assert(kJumpTab % 64 == 0, "jump table in whole 64-entry chunks");
        if (a.s_quads) {
            const uint4* src = a.jump + kJumpCols;
            for (int c = w; c < kJumpTab / 64; c += kUpdThreads / 64)
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + c * 64 + lane),
                                                 (__attribute__((address_space(3))) void*)&jtab[c * 64], 16, 0, 0);
        }
    }
    M<R, N> S;
    if (active) S = staple_nu<R, N, NREAL, TW>(Ts, a.g, s, i, p, mu, nu, a.tw);
    else m_zero(S);
    // Wavefront 0 issues the load of the link it updates before the
    // barrier, so its latency overlaps the wait for the other shares rather
    // than following it.
    M<R, N> U;
    if (w == 0 && active) U = Tb.load_u(p, mu, i);
    if (w > 0) {
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int c = 0; c < N; ++c) {
                lds[w - 1][2 * (r * N + c)][lane] = S.e[r][c].re;
                lds[w - 1][2 * (r * N + c) + 1][lane] = S.e[r][c].im;
            }
    }
    if constexpr (RK == RK_XOR128) __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the table loads landed
    __syncthreads();
    if (w > 0) {
        if constexpr (kStream) {
            if (a.s_quads) {
                const int half = w - 1;
                uint2 o;
                if constexpr (RK == RK_XOR128) {
                    o = xor128_jump_tab_half(rs, jtab, half);
                } else {
                    o = half ? make_uint2(pm_jump(rs.z, a.pm_ak), pm_jump(rs.w, a.pm_ak))
                             : make_uint2(pm_jump(rs.x, a.pm_ak), pm_jump(rs.y, a.pm_ak));
                }
                reinterpret_cast<uint2*>(a.rng_out + i)[half] = o;
            }
        }
        return;
    }

    M<R, N> S1, S2;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
            S1.e[r][c] = {lds[0][2 * (r * N + c)][lane], lds[0][2 * (r * N + c) + 1][lane]};
            S2.e[r][c] = {lds[1][2 * (r * N + c)][lane], lds[1][2 * (r * N + c) + 1][lane]};
        }
    if constexpr (MEAS) {
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int c = 0; c < N; ++c) {
                lds[2][2 * (r * N + c)][lane] = S.e[r][c].re;
                lds[2][2 * (r * N + c) + 1][lane] = S.e[r][c].im;
            }
    }
    M<R, N> A = S;
    m_acc(A, S1);
    m_acc(A, S2);

    double trace = 0.0;
    if (active) {
        M<R, N> Alow;   // unused by the update itself
        (void)update_link<R, N, NREAL, OVER, RK, false, TW>(a, Tb, i, s, A, Alow, rs, U);
        if constexpr (MEAS) {
            // ReTr(U' sum_{nu<mu} S_nu): directions below mu are shares
            // 0 .. mu-1 (nu = w for w < mu), summed in that order; the
            // shares come back from LDS (this lane's own slots)
            auto share = [&](int k) {
                M<R, N> X;
#pragma unroll
                for (int r = 0; r < N; ++r)
#pragma unroll
                    for (int c = 0; c < N; ++c) X.e[r][c] = {lds[k][2 * (r * N + c)][lane], lds[k][2 * (r * N + c) + 1][lane]};
                return X;
            };
            m_zero(Alow);
            if (mu > 0) m_acc(Alow, share(2));
            if (mu > 1) m_acc(Alow, share(0));
            if (mu > 2) m_acc(Alow, share(1));
            trace = m_retr_mul(U, Alow);
        }
    }
    if constexpr (MEAS) {
        trace = wave_sum(trace);
        // slot = 64-link unit, so the final sum runs in lattice order
        // whatever the launch order or kernel shape
        if (lane == 0) a.partials[(size_t)(a.meas_mode - 1) * nunits + blk] = trace;
    }
}

template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS, bool TW = false, int MU = -1>
__global__ __launch_bounds__(kUpdThreads) void k_update(UpdArgs a)
{
    update3<R, N, NREAL, OVER, RK, MEAS, TW, MU>(a, upd_block(a), gridDim.x);
}

// One wavefront per 64 links: each wavefront gathers and sums the three
// staple shares of its own links (one batch of gathers per direction) and
// updates them; a workgroup is 4 independent wavefronts.  More links are in
// flight per CU than with kUpdThreads workgroups, at the cost of a longer
// path per wavefront.  The 64-link unit map is the same as upd_block's.
constexpr int kWaves1 = 4;
__device__ __forceinline__ uint32_t upd_unit1(const UpdArgs& a, int w)
{
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    if (a.xcd_map == 0 || (nb % 8u)) return b * kWaves1 + (uint32_t)w;
    const uint32_t per = nb * kWaves1 / 8u, k = b % 8u, l = (b / 8u) * kWaves1 + (uint32_t)w;
    if (a.xcd_map >= 2) return block_unit(a, k, l);
    return k * per + l;
}

// Body of the one-wavefront update for 64-link unit `unit` of `units`.
template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS>
__device__ __forceinline__ void update1(const UpdArgs& a, uint32_t unit, uint32_t units)
{
    const int lane = threadIdx.x & 63;
    const uint32_t i = (a.unit0 + unit) * kLinks + lane;   // < number of streams
    const bool active = i < a.g.vh;
    Table<R, N, NREAL> Tb{(typename Vec2<R>::T*)a.table, a.g.vh_pad};
    const Table<R, N, NREAL> Ts{(typename Vec2<R>::T*)a.src, a.g.vh_pad};
    const int p = a.parity, mu = a.mu;
    const Site s = site_of(a.g, active ? i : 0u, p);

    constexpr bool kStream = (RK == RK_XOR128 || RK == RK_PM);
    uint4 rs = make_uint4(0, 0, 0, 0);
    if constexpr (kStream) {
        if (active || a.s_quads) rs = a.rng[i];
    }
    M<R, N> A, Alow;
    m_zero(A);
    m_zero(Alow);
    if (active) {
        for (int k = 0; k < 3; ++k) {
            const int nu = k + (k >= mu ? 1 : 0);
            const M<R, N> S = staple_nu<R, N, NREAL>(Ts, a.g, s, i, p, mu, nu);
            m_acc(A, S);
            if constexpr (MEAS) {
                if (nu < mu) m_acc(Alow, S);
            }
        }
    }
    double trace = 0.0;
    if (active) {
        M<R, N> U = Tb.load_u(p, mu, i);
        trace = update_link<R, N, NREAL, OVER, RK, MEAS>(a, Tb, i, s, A, Alow, rs, U);
    }
    if constexpr (kStream) {
        // Reference stream layout: every stream advances by exactly s_quads
        // quads from its start state.
        if (a.s_quads) {
            uint4 o;
            if constexpr (RK == RK_XOR128) {
                const uint2 lo = xor128_jump_half(rs, a.jump, 0), hi = xor128_jump_half(rs, a.jump, 1);
                o = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                o = make_uint4(pm_jump(rs.x, a.pm_ak), pm_jump(rs.y, a.pm_ak), pm_jump(rs.z, a.pm_ak),
                               pm_jump(rs.w, a.pm_ak));
            }
            a.rng_out[i] = o;
        }
    }
    if constexpr (MEAS) {
        trace = wave_sum(trace);
        if (lane == 0) a.partials[(size_t)(a.meas_mode - 1) * units + unit] = trace;
    }
}

template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS>
__global__ __launch_bounds__(kWaves1 * kLinks) void k_update1(UpdArgs a, uint32_t units)
{
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t unit = upd_unit1(a, w);
    if (unit >= units) return;   // whole wavefront: beyond the last 64-link unit
    update1<R, N, NREAL, OVER, RK, MEAS>(a, unit, units);
}

// ---------------------------------------------------------------- PRNG kernels
// The reference performs one produce() at initialisation (random.cpp:276):
// advance every stream by `quads` quads.
__global__ __launch_bounds__(kBlock) void k_xor128_skip(uint4* st, uint32_t n, uint32_t quads)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint4 s = st[i];
    for (uint32_t k = 0; k < 4 * quads; ++k) (void)xor128_step(s);
    st[i] = s;
}

__global__ __launch_bounds__(kBlock) void k_pm_skip(uint4* st, uint32_t n, uint32_t quads)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint4 s = st[i];
    for (uint32_t k = 0; k < quads; ++k) (void)pm_quad(s);
    st[i] = s;
}

// Replay of many heatbath half-steps in one pass: apply a precomputed jump
// (XOR128: GF(2) matrix columns; PM: multiplier) to every stream.
__global__ __launch_bounds__(kBlock) void k_xor128_jump_apply(uint4* st, uint32_t n, const uint4* cols)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint4 s = st[i];
    const uint2 lo = xor128_jump_half(s, cols, 0), hi = xor128_jump_half(s, cols, 1);
    st[i] = make_uint4(lo.x, lo.y, hi.x, hi.y);
}

__global__ __launch_bounds__(kBlock) void k_pm_jump_apply(uint4* st, uint32_t n, uint32_t ak)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint4 s = st[i];
    st[i] = make_uint4(pm_jump(s.x, ak), pm_jump(s.y, ak), pm_jump(s.z, ak), pm_jump(s.w, ak));
}

// Reference produce kernels (buffer semantics): quads[i * n + g].
__global__ __launch_bounds__(kBlock) void k_xor128_produce(uint4* st, uint32_t n, uint32_t samples, float4* quads)
{
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    uint4 s = st[g];
    for (uint32_t i = 0; i < samples; ++i) quads[(size_t)i * n + g] = xor128_quad(s);
    st[g] = s;
}

__global__ __launch_bounds__(kBlock) void k_pm_produce(uint4* st, uint32_t n, uint32_t samples, float4* quads)
{
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    uint4 s = st[g];
    for (uint32_t i = 0; i < samples; ++i) quads[(size_t)i * n + g] = pm_quad(s);
    st[g] = s;
}

// Reference RANLUX produce() (prngcl_ranlux.cl:534-563): every stream draws
// `samples` quads, written to quads[k * n + g] unless quads is null (the
// initial produce of PRNG::initialize only advances the streams).  state:
// 7 float4 per stream at state[q * n + g]; q = 0..5 hold ring values
// s[4q .. 4q+3], q = 6 = (23, 9, quad counter of the block, carry).  All
// streams advance together, so the quad counter is uniform.
template <int LUX>
__device__ __forceinline__ void ranlux_body(float4* state, uint32_t n, uint32_t samples, float4* quads, uint32_t g)
{
    // the ring as 24-bit integers (the stored floats are k 2^-24 exactly)
    uint32_t s[24];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float4 v = state[(size_t)q * n + g];
        s[4 * q] = (uint32_t)(v.x * 16777216.0f);
        s[4 * q + 1] = (uint32_t)(v.y * 16777216.0f);
        s[4 * q + 2] = (uint32_t)(v.z * 16777216.0f);
        s[4 * q + 3] = (uint32_t)(v.w * 16777216.0f);
    }
    const float4 c6 = state[(size_t)6 * n + g];
    int k = __builtin_amdgcn_readfirstlane((int)c6.z);
    uint32_t carry = c6.w > 0.0f ? 1u : 0u;
    for (uint32_t m = 0; m < samples; ++m) {
        if (k >= 6) {
            rl_skip<RlSkip<LUX>::full, RlSkip<LUX>::rem>(s, carry);
            k = 0;
        }
        float4 q;
        switch (k) {
            case 0: q = rl_quad<0>(s, carry); break;
            case 1: q = rl_quad<1>(s, carry); break;
            case 2: q = rl_quad<2>(s, carry); break;
            case 3: q = rl_quad<3>(s, carry); break;
            case 4: q = rl_quad<4>(s, carry); break;
            default: q = rl_quad<5>(s, carry); break;
        }
        if (quads) quads[(size_t)m * n + g] = q;
        ++k;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q)
        state[(size_t)q * n + g] = make_float4((float)s[4 * q] * kRlTwom24, (float)s[4 * q + 1] * kRlTwom24,
                                               (float)s[4 * q + 2] * kRlTwom24, (float)s[4 * q + 3] * kRlTwom24);
    state[(size_t)6 * n + g] = make_float4(23.0f, 9.0f, (float)k, carry ? kRlTwom24 : 0.0f);
}

template <int LUX>
__global__ __launch_bounds__(kBlock) void k_ranlux_produce(float4* state, uint32_t n, uint32_t samples, float4* quads)
{
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    ranlux_body<LUX>(state, n, samples, quads, g);
}

void launch_ranlux_produce(int lux, float4* state, uint32_t n, uint32_t samples, float4* quads, hipStream_t st)
{
    const uint32_t nb = (n + kBlock - 1) / kBlock;
    switch (lux) {
        case 0: hipLaunchKernelGGL(k_ranlux_produce<0>, dim3(nb), dim3(kBlock), 0, st, state, n, samples, quads); break;
        case 1: hipLaunchKernelGGL(k_ranlux_produce<1>, dim3(nb), dim3(kBlock), 0, st, state, n, samples, quads); break;
        case 2: hipLaunchKernelGGL(k_ranlux_produce<2>, dim3(nb), dim3(kBlock), 0, st, state, n, samples, quads); break;
        case 3: hipLaunchKernelGGL(k_ranlux_produce<3>, dim3(nb), dim3(kBlock), 0, st, state, n, samples, quads); break;
        default: hipLaunchKernelGGL(k_ranlux_produce<4>, dim3(nb), dim3(kBlock), 0, st, state, n, samples, quads); break;
    }
}

// Host seeding of the RANLUX table as PRNG::initialize / RL_initialize
// (random.cpp:425-520): instance 0 from rand() % (2^31 - 1), the others from
// rand(); 24 values from the L'Ecuyer MLCG 40014/53668/12211
// (prngcl_ranlux.cl:98-135), carry = 2^-24 if the last value is 0; the block
// counter starts at 6 so the first quad is preceded by a skip.
void ranlux_seed_table(unsigned n, std::vector<float>& tab)
{
    tab.assign((size_t)28 * n, 0.0f);
    for (unsigned g = 0; g < n; ++g) {
        int j = (g == 0) ? rand() % 2147483647 : rand();
        float v[24];
        for (int q = 0; q < 24; ++q) {
            const int kk = j / 53668;
            j = 40014 * (j - kk * 53668) - kk * 12211;
            if (j < 0) j += 2147483563;
            v[q] = (float)(j % 16777216) * 5.9604644775390625e-08f;
        }
        for (int q = 0; q < 6; ++q)
            for (int c = 0; c < 4; ++c) tab[4 * ((size_t)q * n + g) + c] = v[4 * q + c];
        float* car = &tab[4 * ((size_t)6 * n + g)];
        car[0] = 23.0f;
        car[1] = 9.0f;
        car[2] = 6.0f;
        car[3] = (v[23] == 0.0f) ? 5.9604644775390625e-08f : 0.0f;
    }
}
// ---------------------------------------------------------------- init kernels
template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_init_cold(void* table, Geom g)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    M<R, N> u;
    m_unit(u);
    for (int p = 0; p < 2; ++p)
        for (int mu = 0; mu < 4; ++mu) T.store_u(p, mu, i, u);
}

// Haar-random start (the reference's lattice_init_hot_* is not Haar,
// su2_update_cl.cl:110-146).  Thread i consumes stream i for the links of its
// two sites in the order parity 0, 1, direction 0..3: one quad per SU(2)
// link, three per SU(3) link.  Box-Muller Gaussians in double
// (t = 1 - u1, 2^-33 if u1 = 1): SU(2) links are a normalised 4-vector
// (uniform on S^3), SU(3) links two complex Gaussian rows through
// Gram-Schmidt with row 3 = conj(row 1 x row 2), a uniform orthonormal
// 2-frame completed to SU(3), which is Haar.
__device__ __forceinline__ void box_muller(float u1, float u2, double& g0, double& g1)
{
    double t = 1.0 - (double)u1;
    if (t <= 0.0) t = 1.1641532182693481e-10;   // 2^-33
    const double r = sqrt(-2.0 * log(t));
    double sn, cs;
    sincos(6.2831853071795864769 * (double)u2, &sn, &cs);
    g0 = r * cs;
    g1 = r * sn;
}

template <typename R, int N, int NREAL, int RK>
__global__ __launch_bounds__(kBlock) void k_init_hot(void* table, Geom g, uint4* rng, const float4* buf, uint32_t bstride)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    constexpr int Q = (N == 3) ? 3 : 1;
    uint4 st = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (RK != RK_BUF) st = rng[i];
    uint32_t kq = 0;   // RK_BUF: next quad of stream i
    for (int p = 0; p < 2; ++p)
        for (int d = 0; d < 4; ++d) {
            double gs[4 * Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                float4 u;
                if constexpr (RK == RK_XOR128) u = xor128_quad(st);
                else if constexpr (RK == RK_PM) u = pm_quad(st);
                else u = buf[(size_t)(kq++) * bstride + i];
                box_muller(u.x, u.y, gs[4 * q], gs[4 * q + 1]);
                box_muller(u.z, u.w, gs[4 * q + 2], gs[4 * q + 3]);
            }
            M<double, N> a;
            if constexpr (N == 3) {
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) a.e[r][c] = {gs[6 * r + 2 * c], gs[6 * r + 2 * c + 1]};
                su3_gram_schmidt(a);
            } else {
                a.e[0][0] = {gs[0], gs[1]};
                a.e[0][1] = {gs[2], gs[3]};
                su2_normalize(a);
            }
            T.store_u(p, d, i, m_cast<R>(a));
        }
    if constexpr (RK != RK_BUF) rng[i] = st;
}

// One SU(2) element from three uniform numbers by M. Di Pierro's map
// (lattice_random2, su2_update_cl.cl:110-146): alpha = pi u, phi = 2 pi v,
// cos theta = 2 w - 1, quaternion a = (cos alpha, sin alpha sin theta cos phi,
// sin alpha sin theta sin phi, sin alpha cos theta), as the matrix
// [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]].  In the link precision.
template <typename R> __device__ __forceinline__ M<R, 2> di_pierro(float u, float v, float w)
{
    const R alpha = R(3.1415926535897932384626433832795) * (R)u;
    const R phi = R(6.2831853071795864769252867665590) * (R)v;
    const R costh = R(2) * (R)w - R(1);
    R sa, ca, sp, cp;
    if constexpr (sizeof(R) == 4) {
        sincosf(alpha, &sa, &ca);
        sincosf(phi, &sp, &cp);
    } else {
        sincos(alpha, &sa, &ca);
        sincos(phi, &sp, &cp);
    }
    const R t = sa * sqrt(R(1) - costh * costh);
    const R a1 = t * cp, a2 = t * sp, a3 = sa * costh;
    M<R, 2> m;
    m.e[0][0] = {ca, a3};
    m.e[0][1] = {a2, a1};
    m.e[1][0] = {-a2, a1};
    m.e[1][1] = {ca, -a3};
    return m;
}

// w embedded in rows / columns p < q of the SU(3) unit matrix.
template <typename R> __device__ __forceinline__ M<R, 3> embed3(const M<R, 2>& w, int p, int q)
{
    M<R, 3> m;
    m_unit(m);
    m.e[p][p] = w.e[0][0];
    m.e[p][q] = w.e[0][1];
    m.e[q][p] = w.e[1][0];
    m.e[q][q] = w.e[1][1];
    return m;
}

// The reference's random ("hot") start, one direction d per launch
// (lattice_init_hot_X..T, suncl.cl:56-161): buf is the flat quad buffer of
// one produce(); the link of site s (lexicographic index y + N2 (z + N3
// (t + N4 x)), the reference's table index) reads quad s (SU(2)) or quads
// s + k rowsize, k = 0..2 (SU(3)).  SU(3) (lattice_random3,
// su3_update_cl.cl:165-225): the SU(2) elements A, B, C of components x, y, z
// (u from the first quad, v from the second, w from the third) in the product
// E12(A) E13(B) E23(C) of subgroup embeddings.
template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_init_random(void* table, Geom g, const float4* buf, uint32_t rowsize,
                                                        int d)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    for (int p = 0; p < 2; ++p) {
        const Site s = site_of(g, i, p);
        const uint32_t gid = (uint32_t)s.y + (uint32_t)g.n2 * ((uint32_t)s.z + (uint32_t)g.n3 *
                                                                 ((uint32_t)s.t + (uint32_t)g.n4 * (uint32_t)s.x));
        if constexpr (N == 2) {
            const float4 q = buf[gid];
            T.store_u(p, d, i, di_pierro<R>(q.x, q.y, q.z));
        } else {
            const float4 q0 = buf[gid], q1 = buf[gid + rowsize], q2 = buf[gid + 2u * rowsize];
            const M<R, 3> a = embed3(di_pierro<R>(q0.x, q1.x, q2.x), 0, 1);
            const M<R, 3> b = embed3(di_pierro<R>(q0.y, q1.y, q2.y), 0, 2);
            const M<R, 3> c = embed3(di_pierro<R>(q0.z, q1.z, q2.z), 1, 2);
            T.store_u(p, d, i, m_mul(m_mul(a, b), c));
        }
    }
}

// lattice_init_gid (suncl.cl:190-347) / lattice_matrixGID
// (suncpp/su3/algebra_su3.cpp:64-80, su2/algebra_su2.cpp:54-64).
template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_init_gid(void* table, Geom g, uint32_t sites)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    for (int p = 0; p < 2; ++p) {
        Site s = site_of(g, i, p);
        double gid = (double)((uint32_t)s.y + (uint32_t)g.n2 * ((uint32_t)s.z + (uint32_t)g.n3 *
                                                                  ((uint32_t)s.t + (uint32_t)g.n4 * (uint32_t)s.x)));
        double fs = (double)sites;
        for (int d = 0; d < 4; ++d) {
            double f = (d + 1) / 1000.0;
            if constexpr (N == 3) {
                M<double, 3> a;
                a.e[0][0] = {sin((f + 10.0 / fs) * gid) / 2.0, cos((f - 10.0 / fs) * gid) / 3.0};
                a.e[0][1] = {sin((f - 18.0 / fs) * gid) / 4.0, cos((f + 12.0 / fs) * gid) / 4.0};
                a.e[0][2] = {sin((f - 12.0 / fs) * gid) / 3.0, cos((f - 15.0 / fs) * gid) / 2.0};
                a.e[1][0] = {sin((f + 15.0 / fs) * gid) / 3.0, cos((f - 11.0 / fs) * gid) / 4.0};
                a.e[1][1] = {sin((f - 11.0 / fs) * gid) / 3.0, cos((f + 16.0 / fs) * gid) / 2.0};
                a.e[1][2] = {sin((f - 16.0 / fs) * gid) / 3.0, cos((f - 18.0 / fs) * gid) / 2.0};
                su3_gram_schmidt(a);
                T.store_u(p, d, i, m_cast<R>(a));
            } else {
                M<double, 2> a;
                double u1r = sin((f + 10.0 / fs) * gid) / 2.0;
                double u2i = cos((f - 10.0 / fs) * gid) / 3.0;
                double u2r = sin((f - 18.0 / fs) * gid) / 4.0;
                double u1i = sqrt(1.0 - u1r * u1r - u2i * u2i - u2r * u2r);
                a.e[0][0] = {u1r, u1i};
                a.e[0][1] = {u2r, u2i};
                a.e[1][0] = {-u2r, u2i};
                a.e[1][1] = {u1r, -u1i};
                T.store_u(p, d, i, m_cast<R>(a));
            }
        }
    }
}

template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_reunitarize(void* table, Geom g)
{
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    for (int p = 0; p < 2; ++p)
        for (int mu = 0; mu < 4; ++mu) {
            M<double, N> a = m_cast<double>(T.load_u(p, mu, i));
            if constexpr (N == 3) su3_gram_schmidt(a);
            else su2_normalize(a);
            T.store_u(p, mu, i, m_cast<R>(a));
        }
}

// ---------------------------------------------------------------- reductions
// Workgroup partial of K doubles per thread: wave shuffles, then LDS, written
// to partials[block * K + k].  Deterministic; no atomics.
// Writes the workgroup's K sums to partials[slot * K + k].
template <int K>
__device__ __forceinline__ void block_partial(double (&v)[K], double* partials, uint32_t slot)
{
    __shared__ double sm[K][kBlock / 64];
    const uint32_t bid = slot;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double x = wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double x = sm[threadIdx.x][0];
        for (int w = 1; w < kBlock / 64; ++w) x += sm[threadIdx.x][w];
        partials[(size_t)bid * K + threadIdx.x] = x;
    }
}

// Final sums of one measurement cycle, one workgroup: thread t adds slots
// t, t + 256, ... in order, then a fixed tree.  Writes one record and
// advances the record index, so replays from a graph need no memory clears.
struct FinArgs {
    // plaquette from the fused even Y/Z/T half-steps: 3 arrays of n_fused
    const double* fused;
    uint32_t n_fused;
    // plaquette from k_plaquette: n_plq slots of 2
    const double* plq;
    uint32_t n_plq;
    // Polyakov loop: n_poly slots of 4
    const double* poly;
    uint32_t n_poly;
    // Wilson loop: n_wil slots of 1
    const double* wil;
    uint32_t n_wil;
    double* plq_rec;
    double* poly_rec;
    double* wil_rec;
    int* rec_index;   // [0] plaquette, [1] polyakov, [2] Wilson loop, [3] Polyakov slices
    int capacity;
    int slices = 0;   // k_slices wrote record rec_index[3]: advance it
    double* a_rec = nullptr;   // action slice records beside the plaquette record
    int a_n = 0;               // N1 + N2 + N3
    int aslices = 0;           // k_slices wrote this plaquette record's action slices
    // F_munu (k_fmunu): n_plq slots of 12, recorded beside the plaquette
    const double* fmunu = nullptr;
    double* f_rec = nullptr;
};

__device__ __forceinline__ double block_sum_1(double v)
{
    __shared__ double sm[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) sm[wave] = v;
    __syncthreads();
    double x = sm[0];
    for (int w = 1; w < kBlock / 64; ++w) x += sm[w];
    return x;
}

// block_sum_1 of K values at once: the same wavefront sums and the same
// order over wavefronts, one pair of barriers for all of them.
template <int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K])
{
    __shared__ double sm[K][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sm[k][wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double x = sm[k][0];
        for (int w = 1; w < kBlock / 64; ++w) x += sm[k][w];
        v[k] = x;
    }
}

__device__ __forceinline__ void finalize_body(const FinArgs& f)
{
    // Every thread's partial sums of the plaquette, Polyakov and Wilson
    // slots first (their loads all go out before any reduction), then one
    // block reduction of the seven values.
    const uint32_t t = threadIdx.x;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // sp, tm, poly[4], wilson
    if (f.fused) {
#pragma unroll 8
        for (uint32_t j = t; j < f.n_fused; j += kBlock) {
            v[0] += f.fused[j];
            v[0] += f.fused[f.n_fused + j];
            v[1] += f.fused[2 * f.n_fused + j];
        }
    } else if (f.plq) {
#pragma unroll 8
        for (uint32_t j = t; j < f.n_plq; j += kBlock) {
            v[0] += f.plq[2 * j];
            v[1] += f.plq[2 * j + 1];
        }
    }
    if (f.poly) {
#pragma unroll 8
        for (uint32_t j = t; j < f.n_poly; j += kBlock)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[2 + k] += f.poly[4 * j + k];
    }
    if (f.wil)
        for (uint32_t j = t; j < f.n_wil; j += kBlock) v[6] += f.wil[j];
    // every thread reads the plaquette record index before the barriers;
    // thread 0 advances the indices after them
    const int rec = f.rec_index[0];
    block_sum_k<7>(v);
    if (f.fused || f.plq) {
        double fm[12];
        if (f.fmunu) {
            for (int k = 0; k < 12; ++k) {
                double a = 0.0;
                for (uint32_t j = t; j < f.n_plq; j += kBlock) a += f.fmunu[12 * j + k];
                fm[k] = block_sum_1(a);
            }
        }
        if (rec < f.capacity) {
            if (t == 0) {
                f.plq_rec[2 * rec] = v[0];
                f.plq_rec[2 * rec + 1] = v[1];
                for (int k = 0; k < 12; ++k) f.f_rec[12 * rec + k] = f.fmunu ? fm[k] : 0.0;
            }
            if (f.a_rec && !f.aslices)   // a plaquette record without action slices
                for (int k = (int)t; k < 2 * f.a_n; k += kBlock) f.a_rec[(size_t)2 * f.a_n * rec + k] = 0.0;
        }
        if (t == 0) f.rec_index[0] = rec + 1;
    }
    if (f.poly && t == 0) {
        const int r = f.rec_index[1];
        if (r < f.capacity)
#pragma unroll
            for (int k = 0; k < 4; ++k) f.poly_rec[4 * r + k] = v[2 + k];
        f.rec_index[1] = r + 1;
    }
    if (f.wil && t == 0) {
        const int r = f.rec_index[2];
        if (r < f.capacity) f.wil_rec[r] = v[6];
        f.rec_index[2] = r + 1;
    }
    if (f.slices && t == 0) f.rec_index[3] += 1;
}

__global__ __launch_bounds__(kBlock) void k_finalize(FinArgs f) { finalize_body(f); }

// Plaquette: per site sum of ReTr U_p over the spatial planes (x,y), (x,z),
// (y,z) and the temporal planes (x,t), (y,t), (z,t), products in FP64
// (lattice_retrace_plaquette3, su3_measurements_cl.cl:527-549).
template <typename R, int N, int NREAL, bool TW = false>
__global__ __launch_bounds__(kBlock) void k_plaquette(void* table, Geom g, double* partials, Twist tw, double* site_pl)
{
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int p = blockIdx.y, q = p ^ 1;
    double v[2] = {0.0, 0.0};
    if (i < g.vh) {
        Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
        Site s = site_of(g, i, p);
        auto ld = [&](int pp, int d, uint32_t idx) {
            M<R, N> m = T.load_u(pp, d, idx);
            if constexpr (TW) tbc(m, tw, d, idx);
            return m_cast<double>(m);
        };
        // the outer loop stays rolled: unrolled, the FP64 matrices of all six
        // planes are live at once (256 VGPRs, spills in the twisted form)
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {
            M<double, N> Ua = ld(p, a, i);
            const uint32_t i_pa = nb_index(g, s, a, +1);
#pragma unroll
            for (int b = a + 1; b < 4; ++b) {
                const uint32_t i_pb = nb_index(g, s, b, +1);
                M<double, N> Ub = ld(p, b, i);
                M<double, N> Ubxa = ld(q, b, i_pa);
                M<double, N> Uaxb = ld(q, a, i_pb);
                // ReTr[Ua(x) Ub(x+a) Ua(x+b)^+ Ub(x)^+] = ReTr[(Ua Ubxa)(Ub Uaxb)^+]
                double tr = m_retr_nd(m_mul(Ua, Ubxa), m_mul(Ub, Uaxb));
                if (b == 3) v[1] += tr;
                else v[0] += tr;
            }
        }
        if (site_pl) {   // this site's spatial and temporal sums, site order y, z, t, x (x slowest)
            const uint32_t j = (uint32_t)s.y + (uint32_t)g.n2 * ((uint32_t)s.z + (uint32_t)g.n3 *
                                                                 ((uint32_t)s.t + (uint32_t)g.n4 * (uint32_t)s.x));
            site_pl[2 * j] = v[0];
            site_pl[2 * j + 1] = v[1];
        }
    }
    block_partial<2>(v, partials, blockIdx.y * gridDim.x + blockIdx.x);
}

// Tr(P G_k) for generator k: Gell-Mann lambda_k (SU(3), k = 1..8) or Pauli
// sigma_k (SU(2), k = 1..3).
template <int N>
__device__ __forceinline__ C<double> gen_trace(const M<double, N>& P, int k)
{
    auto add = [](C<double> a, C<double> b) { return C<double>{a.re + b.re, a.im + b.im}; };
    auto sub = [](C<double> a, C<double> b) { return C<double>{a.re - b.re, a.im - b.im}; };
    auto mul_i = [](C<double> a) { return C<double>{-a.im, a.re}; };
    if constexpr (N == 3) {
        switch (k) {
            case 1: return add(P.e[0][1], P.e[1][0]);
            case 2: return mul_i(sub(P.e[0][1], P.e[1][0]));
            case 3: return sub(P.e[0][0], P.e[1][1]);
            case 4: return add(P.e[0][2], P.e[2][0]);
            case 5: return mul_i(sub(P.e[0][2], P.e[2][0]));
            case 6: return add(P.e[1][2], P.e[2][1]);
            case 7: return mul_i(sub(P.e[1][2], P.e[2][1]));
            default: {
                const double r3 = 0.57735026918962576451;   // 1 / sqrt(3)
                const C<double> d = sub(add(P.e[0][0], P.e[1][1]), add(P.e[2][2], P.e[2][2]));
                return C<double>{d.re * r3, d.im * r3};
            }
        }
    } else {
        switch (k) {
            case 1: return add(P.e[0][1], P.e[1][0]);
            case 2: return mul_i(sub(P.e[0][1], P.e[1][0]));
            default: return sub(P.e[0][0], P.e[1][1]);
        }
    }
}

// F_munu measurement (lattice_measurement_plq with FMUNU / F0MU,
// sun_measurements_cl.cl:710-1173; lattice_retrace_plaquette3_F /
// lattice_retrace_plaquette2_F, su3_measurements_cl.cl:550-590,
// su2_measurements_cl.cl:234-260): for the planes (x,y), (x,z), (y,z) (H
// field) or (x,t), (y,t), (z,t) (E field, efield = 1), the plaquette
// P = U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+ in FP64 and F = -i Tr(P G_k) / 2
// for the generators g1 and g2 (SU(2): imaginary part 0, as the reference).
// Block partials: 12 sums, [generator][plane][re, im].
template <typename R, int N, int NREAL, bool TW = false>
__global__ __launch_bounds__(kBlock) void k_fmunu(void* table, Geom g, double* partials, int efield, int g1, int g2,
                                                  Twist tw)
{
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int p = blockIdx.y, q = p ^ 1;
    double v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = 0.0;
    if (i < g.vh) {
        Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
        const Site s = site_of(g, i, p);
        const int pa[3] = {efield ? 0 : 0, efield ? 1 : 0, efield ? 2 : 1};
        const int pb[3] = {efield ? 3 : 1, efield ? 3 : 2, efield ? 3 : 2};
        for (int pl = 0; pl < 3; ++pl) {
            const int a = pa[pl], b = pb[pl];
            auto ld = [&](int pp, int d, uint32_t idx) {
                M<R, N> m = T.load_u(pp, d, idx);
                if constexpr (TW) tbc(m, tw, d, idx);
                return m_cast<double>(m);
            };
            const M<double, N> Ua = ld(p, a, i);
            const M<double, N> Ub = ld(p, b, i);
            const M<double, N> Ubxa = ld(q, b, nb_index(g, s, a, +1));
            const M<double, N> Uaxb = ld(q, a, nb_index(g, s, b, +1));
            const M<double, N> X = m_mul(Ua, Ubxa), Y = m_mul(Ub, Uaxb);
            M<double, N> P;   // X Y^+
#pragma unroll
            for (int r = 0; r < N; ++r)
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    C<double> acc{0.0, 0.0};
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        const C<double> x = X.e[r][k], y = Y.e[c][k];
                        acc.re += x.re * y.re + x.im * y.im;
                        acc.im += x.im * y.re - x.re * y.im;
                    }
                    P.e[r][c] = acc;
                }
            const int gens[2] = {g1, g2};
            for (int gi = 0; gi < 2; ++gi) {
                const C<double> z = gen_trace<N>(P, gens[gi]);
                v[6 * gi + 2 * pl] += 0.5 * z.im;                       // Re(-i z / 2)
                v[6 * gi + 2 * pl + 1] += (N == 2) ? 0.0 : -0.5 * z.re;  // Im(-i z / 2)
            }
        }
    }
    block_partial<12>(v, partials, blockIdx.y * gridDim.x + blockIdx.x);
}

// Polyakov loop per spatial site, product of the N4 time links in link
// precision and traces in FP64 (polyakov.cl:54-141).  Eight lanes share one
// spatial site: lane l multiplies a contiguous chunk of ceil(N4/8) time links
// and the eight partial products are combined in order by a three-level
// shuffle tree, so a wavefront holds 8 independent loads per site in flight.
template <typename R, int N>
__device__ __forceinline__ M<R, N> shfl_down_m(const M<R, N>& a, int d)
{
    M<R, N> b;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            b.e[i][j].re = __shfl_down(a.e[i][j].re, d, 64);
            b.e[i][j].im = __shfl_down(a.e[i][j].im, d, 64);
        }
    return b;
}

// R x T Wilson loop (wilson_loop.cl:55-465; su3cpu.cpp:1435-1522), one
// thread per site x, summed over the three spatial directions d:
// ReTr(B Rm Tm^+ Lm^+) with B = U_t(x) .. U_t(x+(T-1)t),
// Rm = U_d(x+Tt) .. U_d(x+Tt+(R-1)d), Tm = U_t(x+Rd) .. U_t(x+Rd+(T-1)t),
// Lm = U_d(x) .. U_d(x+(R-1)d).  Products in link precision, traces and
// sums in FP64; one partial per workgroup.
template <typename R, int N, int NREAL, bool TW>
__device__ __forceinline__ M<R, N> line_product(const Table<R, N, NREAL>& T, const Geom& g, const Twist& tw, Site s,
                                                int dir, int len)
{
    auto ld = [&](const Site& c) {
        const uint32_t idx = cindex(g, c.x, c.y, c.z, c.t);
        M<R, N> m = T.load((c.x + c.y + c.z + c.t) & 1, dir, idx);
        if constexpr (TW) tbc(m, tw, dir, idx);
        return m;
    };
    M<R, N> P = ld(s);
    for (int k = 1; k < len; ++k) {
        shift(g, s, dir, +1);
        P = m_mul(P, ld(s));
    }
    return P;
}

template <typename R, int N, int NREAL, bool TW = false>
__global__ __launch_bounds__(kBlock) void k_wilson(void* table, Geom g, int wr, int wt, double* partials, Twist tw)
{
    const uint32_t site = blockIdx.x * kBlock + threadIdx.x;
    double acc = 0.0;
    if (site < 2u * g.vh) {
        Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
        Site x;
        uint32_t r = site;
        x.y = (int)(r % (uint32_t)g.n2);
        r /= (uint32_t)g.n2;
        x.z = (int)(r % (uint32_t)g.n3);
        r /= (uint32_t)g.n3;
        x.t = (int)(r % (uint32_t)g.n4);
        x.x = (int)(r / (uint32_t)g.n4);
        const M<R, N> B = line_product<R, N, NREAL, TW>(T, g, tw, x, 3, wt);
        Site xt = x;
        for (int k = 0; k < wt; ++k) shift(g, xt, 3, +1);
        for (int d = 0; d < 3; ++d) {
            Site xr = x;
            for (int k = 0; k < wr; ++k) shift(g, xr, d, +1);
            const M<R, N> Rm = line_product<R, N, NREAL, TW>(T, g, tw, xt, d, wr);
            const M<R, N> Tm = line_product<R, N, NREAL, TW>(T, g, tw, xr, 3, wt);
            const M<R, N> Lm = line_product<R, N, NREAL, TW>(T, g, tw, x, d, wr);
            acc += m_retr_nd(m_mul_nd(m_mul(B, Rm), Tm), Lm);
        }
    }
    double v[1] = {acc};
    block_partial<1>(v, partials, blockIdx.x);
}

// Polyakov-loop body for workgroup `blk`: 8 threads per spatial site, thread
// k taking t-slot l = k / 32 (the links t in [l c, (l+1) c), c = ceil(N4/8))
// of site j = 32 blk + k % 32, so each wavefront's loads cover 32
// consecutive sites (whole cache lines) at two t-slots.  The ordered product
// over the slots is formed through LDS as ((P0 P1)(P2 P3))((P4 P5)(P6 P7)).
constexpr int kPolySites = kBlock / 8;
template <typename R, int N, int NREAL>
__device__ __forceinline__ void polyakov_body(void* table, const Geom& g, double* partials, uint32_t blk,
                                              double* site_tr = nullptr)
{
    __shared__ R pm[8][2 * N * N][kPolySites];
    const int ks = (int)(threadIdx.x % kPolySites);
    const int l8 = (int)(threadIdx.x / kPolySites);
    const uint32_t j = blk * kPolySites + (uint32_t)ks;
    const uint32_t vs = (uint32_t)g.n1 * g.n2 * g.n3;
    Table<R, N, NREAL> T{(typename Vec2<R>::T*)table, g.vh_pad};
    M<R, N> P;
    m_unit(P);
    if (j < vs) {
        const int y = (int)(j % (uint32_t)g.n2);
        const int z = (int)((j / (uint32_t)g.n2) % (uint32_t)g.n3);
        const int x = (int)(j / ((uint32_t)g.n2 * g.n3));
        const int chunk = (g.n4 + 7) / 8;
        const int t0 = l8 * chunk;
        const int t1 = min(t0 + chunk, g.n4);
        for (int t = t0; t < t1; ++t) {
            const M<R, N> L = T.load((x + y + z + t) & 1, 3, cindex(g, x, y, z, t));
            P = (t == t0) ? L : m_mul(P, L);
        }
    }
    auto put = [&](const M<R, N>& a) {
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int c = 0; c < N; ++c) {
                pm[l8][2 * (r * N + c)][ks] = a.e[r][c].re;
                pm[l8][2 * (r * N + c) + 1][ks] = a.e[r][c].im;
            }
    };
    put(P);
    __syncthreads();
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
        // slot l (a multiple of 2d) takes the product slot l + d formed at
        // the previous level; no slot read here is written here
        if ((l8 & (2 * d - 1)) == 0) {
            M<R, N> Q;
#pragma unroll
            for (int r = 0; r < N; ++r)
#pragma unroll
                for (int c = 0; c < N; ++c)
                    Q.e[r][c] = {pm[l8 + d][2 * (r * N + c)][ks], pm[l8 + d][2 * (r * N + c) + 1][ks]};
            P = m_mul(P, Q);
            if (d < 4) put(P);
        }
        __syncthreads();
    }
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < vs && l8 == 0) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            re += (double)P.e[k][k].re;
            im += (double)P.e[k][k].im;
        }
        const double a2 = re * re + im * im;
        v[0] = re;
        v[1] = im;
        v[2] = a2;
        v[3] = a2 * a2;
        if (site_tr) {   // per spatial site, for the slice sums
            site_tr[2 * j] = re;
            site_tr[2 * j + 1] = im;
        }
    }
    block_partial<4>(v, partials, blk);
}

template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_polyakov(void* table, Geom g, double* partials, double* site_tr)
{
    polyakov_body<R, N, NREAL>(table, g, partials, blockIdx.x, site_tr);
}

// SU(2) FP32 links between the lattice table and unit quaternions
// q[mu * V + site] = (a0, a1, a2, a3), U = [[a0 + i a3, a2 + i a1],
// [-a2 + i a1, a0 - i a3]], site in the canonical order of the C ABI.
// blockIdx.y = parity * 4 + mu, so parity and direction are uniform.
__global__ __launch_bounds__(kBlock) void k_su2_to_quats(void* table, Geom g, float4* q)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    const int p = (int)blockIdx.y >> 2, mu = (int)blockIdx.y & 3;
    Table<float, 2, 4> T{(typename Vec2<float>::T*)table, g.vh_pad};
    const Site s = site_of(g, i, p);
    const M<float, 2> U = T.load_u(p, mu, i);
    const uint32_t site = (uint32_t)s.y + (uint32_t)g.n2 * ((uint32_t)s.z + (uint32_t)g.n3 *
                                                              ((uint32_t)s.t + (uint32_t)g.n4 * (uint32_t)s.x));
    q[(size_t)mu * 2u * g.vh + site] = make_float4(U.e[0][0].re, U.e[0][1].im, U.e[0][1].re, U.e[0][0].im);
}

__global__ __launch_bounds__(kBlock) void k_quats_to_su2(void* table, Geom g, const float4* q)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.vh) return;
    const int p = (int)blockIdx.y >> 2, mu = (int)blockIdx.y & 3;
    Table<float, 2, 4> T{(typename Vec2<float>::T*)table, g.vh_pad};
    const Site s = site_of(g, i, p);
    const uint32_t site = (uint32_t)s.y + (uint32_t)g.n2 * ((uint32_t)s.z + (uint32_t)g.n3 *
                                                              ((uint32_t)s.t + (uint32_t)g.n4 * (uint32_t)s.x));
    const float4 a = q[(size_t)mu * 2u * g.vh + site];
    M<float, 2> U;
    U.e[0][0] = {a.x, a.w};
    U.e[0][1] = {a.z, a.y};
    U.e[1][0] = {-a.z, a.y};
    U.e[1][1] = {a.x, -a.w};
    T.store_u(p, mu, i, U);
}

// Sums per x, y and z slice of a per-site pair of values, in a fixed order:
// workgroup b takes slice x = b (b < N1), y = b - N1, or z = b - N1 - N2 and
// writes record rec_index[slot] at position b.  Sites are indexed
// y + N2 (z + N3 (t + nt x)); nt = 1 for the spatial sites of the Polyakov
// loop (lattice_polyakov_diff_x/y/z + reduce_polyakov_diff_*_double2,
// polyakov.cl:142-560: (sum Re Tr P, sum Im Tr P)), nt = N4 for the
// plaquette sums of the action per slice (lattice_action_diff_x/y/z +
// reduce_action_diff_*_double2, sun_measurements_cl.cl:186-708).  The
// finaliser advances the record index.
__global__ __launch_bounds__(kBlock) void k_slices(const double* site, Geom g, uint32_t nt, double* slice_rec, int cap,
                                                   const int* rec_index, int slot)
{
    __shared__ double red[2][kBlock];
    const uint32_t b = blockIdx.x, n1 = (uint32_t)g.n1, n2 = (uint32_t)g.n2, n3 = (uint32_t)g.n3;
    const int d = b < n1 ? 0 : (b < n1 + n2 ? 1 : 2);
    const uint32_t v = d == 0 ? b : (d == 1 ? b - n1 : b - n1 - n2);
    const uint32_t count = n1 * n2 * n3 * nt / (d == 0 ? n1 : (d == 1 ? n2 : n3));
    double re = 0.0, im = 0.0;
    for (uint32_t k = threadIdx.x; k < count; k += kBlock) {
        uint32_t j;
        if (d == 0) {          // x = v: y, z, t
            j = (k % n2) + n2 * ((k / n2) % n3 + n3 * (k / (n2 * n3) + nt * v));
        } else if (d == 1) {   // y = v: z, t, x
            j = v + n2 * ((k % n3) + n3 * ((k / n3) % nt + nt * (k / (n3 * nt))));
        } else {               // z = v: y, t, x
            j = (k % n2) + n2 * (v + n3 * ((k / n2) % nt + nt * (k / (n2 * nt))));
        }
        re += site[2 * j];
        im += site[2 * j + 1];
    }
    red[0][threadIdx.x] = re;
    red[1][threadIdx.x] = im;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int rec = rec_index[slot];
        if (rec < cap) {
            double* o = slice_rec + 2 * ((size_t)rec * (n1 + n2 + n3) + b);
            o[0] = red[0][0];
            o[1] = red[1][0];
        }
    }
}

// ---------------------------------------------------------------- batched ensembles
// Many independent lattices ("chains") updated by one launch per half-step.
// ChainDesc holds what the kernels need for one chain; a host-built work map
// gives each workgroup (or wavefront) its chain and 64-link unit, with
// chain = ~0u marking idle entries.  The map is built so that each chain's
// units run on one XCD.
struct ChainDesc {
    UpdArgs a;            // parity, mu, meas_mode, rng, rng_out are set per launch
    uint4* rng_buf[2];    // XOR128 / PM states (alternating)
    float4* rl_state;     // RANLUX states
    uint32_t nunits;      // 64-link units = streams / 64
    double* part_poly;    // Polyakov partials of this chain
    FinArgs fin;          // finaliser arguments (sources selected per launch)
};

__device__ __forceinline__ UpdArgs chain_args(const ChainDesc* cd, uint32_t c, int parity, int mu, int meas_mode,
                                              int rng_cur)
{
    UpdArgs a = cd[c].a;
    a.parity = parity;
    a.mu = mu;
    a.meas_mode = meas_mode;
    a.rng = cd[c].rng_buf[rng_cur];
    a.rng_out = cd[c].rng_buf[rng_cur ^ 1];
    return a;
}

template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS>
__global__ __launch_bounds__(kUpdThreads) void k_update_batch3(const ChainDesc* cd, const uint2* map, int parity,
                                                               int mu, int meas_mode, int rng_cur)
{
    const uint32_t c = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const uint32_t u = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    if (c == ~0u) return;   // whole workgroup idle
    const UpdArgs a = chain_args(cd, c, parity, mu, meas_mode, rng_cur);
    update3<R, N, NREAL, OVER, RK, MEAS>(a, u, cd[c].nunits);
}

template <typename R, int N, int NREAL, bool OVER, int RK, bool MEAS>
__global__ __launch_bounds__(kWaves1 * kLinks) void k_update_batch1(const ChainDesc* cd, const uint2* map, int parity,
                                                                    int mu, int meas_mode, int rng_cur)
{
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t e = blockIdx.x * kWaves1 + (uint32_t)w;
    const uint32_t c = __builtin_amdgcn_readfirstlane(map[e].x);
    const uint32_t u = __builtin_amdgcn_readfirstlane(map[e].y);
    if (c == ~0u) return;   // whole wavefront idle
    const UpdArgs a = chain_args(cd, c, parity, mu, meas_mode, rng_cur);
    update1<R, N, NREAL, OVER, RK, MEAS>(a, u, cd[c].nunits);
}

// RANLUX produce for every chain: one wavefront-sized workgroup per unit.
template <int LUX>
__global__ __launch_bounds__(kLinks) void k_ranlux_produce_batch(const ChainDesc* cd, const uint2* map, uint32_t samples)
{
    const uint32_t c = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const uint32_t u = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    if (c == ~0u) return;
    const uint32_t n = cd[c].nunits * kLinks;
    ranlux_body<LUX>(cd[c].rl_state, n, samples, const_cast<float4*>(cd[c].a.prn), u * kLinks + threadIdx.x);
}

template <typename R, int N, int NREAL>
__global__ __launch_bounds__(kBlock) void k_polyakov_batch(const ChainDesc* cd, const uint2* map)
{
    const uint32_t c = __builtin_amdgcn_readfirstlane(map[blockIdx.x].x);
    const uint32_t b = __builtin_amdgcn_readfirstlane(map[blockIdx.x].y);
    if (c == ~0u) return;
    polyakov_body<R, N, NREAL>(cd[c].a.table, cd[c].a.g, cd[c].part_poly, b);
}

// One workgroup per chain.  flags: 1 fused plaquette, 4 Polyakov.
__global__ __launch_bounds__(kBlock) void k_finalize_batch(const ChainDesc* cd, int flags)
{
    FinArgs f = cd[blockIdx.x].fin;
    if (!(flags & 1)) f.fused = nullptr;
    f.plq = nullptr;
    if (!(flags & 4)) f.poly = nullptr;
    f.wil = nullptr;
    finalize_body(f);
}

}  // namespace

// ============================================================================
// Host side
// ============================================================================
__device__ __align__(256) double scratch[kScratchBytesDev / sizeof(double)];

struct qcd_lattice {
    qcd_params p;
    Geom g;
    int np;                // complex planes per link
    size_t elem;           // bytes per complex value
    void* table = nullptr;
    size_t table_bytes = 0;
    uint4* rng_buf[2] = {nullptr, nullptr};   // generator states; rng_buf[rng_cur] is current
    uint4* jump_cols = nullptr;               // XOR128 jump-ahead matrix: 128 columns, then the 4-bit table
    float4* rl_state = nullptr;               // RANLUX: 7 float4 per stream
    float4* prn_buf = nullptr;                // RANLUX: produced quads, prn_quads per stream
    uint32_t prn_quads = 0;
    uint32_t pm_ak = 1;                       // Park-Miller jump multiplier
    int rng_cur = 0;
    uint32_t n_streams = 0;   // PRNGSTEP = align64(V/2): links of a half-step's grid, the PRN buffer's stride
    // Generator instances (PRNG_instances): n_streams, or the count the
    // parameters give.  With quad_buffer, every heatbath half-step is
    // preceded by a produce() of s_quads quads per instance into prn_buf
    // (quad q = sample * n_inst + instance) and link i reads quad
    // i + k * n_streams, the reference's PRN buffer itself.
    uint32_t n_inst = 0;
    bool quad_buffer = false;
    uint32_t s_quads = 0;   // S
    double* partials = nullptr;     // one allocation, three regions:
    double* part_fused = nullptr;   //   3 x nb_fused  (fused plaquette)
    double* part_plq = nullptr;     //   2 x plq blocks (k_plaquette)
    double* part_poly = nullptr;    //   4 x poly_blocks (k_polyakov)
    uint32_t nb_fused = 0;
    int* rec_index = nullptr;       // [0] plaquette, [1] polyakov
    double* plq_rec = nullptr;
    double* poly_rec = nullptr;
    double* wil_rec = nullptr;          // Wilson loop records
    double* part_wil = nullptr;         //   wil_blocks (k_wilson)
    double* site_tr = nullptr;          // Tr P per spatial site (2 doubles), for slice sums
    double* slice_rec = nullptr;        // Polyakov slice records: [rec][N1 + N2 + N3][2]
    bool want_slices = false;           // the Polyakov launch being enqueued feeds the slice sums
    double* site_pl = nullptr;          // per-site plaquette sums (2 doubles), allocated on first use
    double* a_rec = nullptr;            // action slice records: [rec][N1 + N2 + N3][spatial, temporal]
    bool want_aslices = false;          // the plaquette launch being enqueued feeds the action slices
    double* part_f = nullptr;           // F_munu block partials: 2 x upd_blocks slots of 12
    double* f_rec = nullptr;            // F_munu records: [rec][generator][plane][re, im]
    int f_gen[2] = {3, 8};              // generators (SU(2): 3 and 2)
    int f_efield = 0;                   // F0MU: planes (x,t), (y,t), (z,t)
    double tw_phi = 0.0, tw_omega = 0.0;   // PHI, OMEGA of the twisted boundary
    Twist tw;                              // what the kernels apply, derived from them
    uint32_t wil_blocks = 0;
    int wil_R = 7, wil_T = 2;           // wilson_R, wilson_T defaults of QCDGPU.cpp
    uint32_t upd_blocks = 0, poly_blocks = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cyc_hb = 0, cyc_or = 0, cyc_mask = 0;
    // cyc_k consecutive cycles recorded as one graph, each cycle's
    // measurement on a second stream beside the next cycle's first three
    // half-steps (QCD_TUNE_CYCLES_PER_GRAPH)
    hipGraph_t graph_k = nullptr;
    hipGraphExec_t exec_k = nullptr;
    int cyc_k = 0;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool rng_ready = false;
    bool hot_run = false;       // produce() sized as in the reference's hot runs (QCD_START_RANDOM)
    bool gid_updates = false;   // INTS = 2: GID known-answer numbers in the heatbath
    // PRNG::produce calls the reference would have made so far
    // (PRNG::PRNG_counter, random.cpp:283): 1 for the initial produce, 4 for
    // a hot start, 1 per heatbath half-step.  Resume replays the streams to
    // this position (suncl.cpp:5349-5351).
    uint32_t prng_counter = 0;
    uint32_t cyc_prn = 0;       // productions in one replay of the captured cycle
    int tune_xcd_map = 2;       // QCD_TUNE_XCD_MAP
    int tune_blocks_per_cu = 0; // QCD_TUNE_BLOCKS_PER_CU (0 = no cap)
    int tune_update_shape = 3;  // QCD_TUNE_UPDATE_SHAPE: wavefronts per 64 links (3 or 1)
    void* shadow = nullptr;     // QCD_TUNE_SHADOW: the copy the update gathers read (race check)
    int tune_cycles = 1;        // QCD_TUNE_CYCLES_PER_GRAPH (1: single-cycle graphs, the faster on MI355X)
    // A chain of a qcd_batch: the batch's device descriptors hold this
    // lattice's buffers and stream sizing, so neither may change (no
    // re-sizing start, no destroy except through the batch).
    bool batch_owned = false;
    // x-slab (one GPU's share of a lattice cut along x): the slices x = 0 and
    // x = N1 - 1 are halos that the updates read and never write; they are
    // refreshed through qcd_slab_face.
    bool slab = false;
    // Kernel timing (GPU::print_time_detailed, clinterface.cpp:1626-1666):
    // with profiling on, every kernel family's launches are timed with a
    // pair of events on the lattice stream and waited for, as the
    // reference's profiling queue serialises them.
    struct ProfStat {
        std::string name;
        unsigned n = 0;
        double sum = 0.0, sumsq = 0.0;   // ms
    };
    bool profiling = false;
    std::vector<ProfStat> prof;
    hipEvent_t prof_ev[2] = {nullptr, nullptr};
    // Kernels created by name (GPU::kernel_init): registry index -> entry of
    // the name table, and how often each ran (GPU_kernels[i].kernel_number_of_starts).
    std::vector<std::pair<int, unsigned>> kernels;
    bool graph_open = false;        // qcd_graph_begin .. qcd_graph_end
    uint32_t graph_counter0 = 0;
    int graph_rng0 = 0;
};

namespace {
// Drops the captured cycle graphs (the single cycle and the multi-cycle one).
void drop_cycle_graphs(qcd_lattice* L)
{
    if (L->exec) (void)hipGraphExecDestroy(L->exec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    if (L->exec_k) (void)hipGraphExecDestroy(L->exec_k);
    if (L->graph_k) (void)hipGraphDestroy(L->graph_k);
    L->exec = nullptr;
    L->graph = nullptr;
    L->exec_k = nullptr;
    L->graph_k = nullptr;
    L->cyc_k = 0;
}
}  // namespace

namespace {

thread_local char g_err[512] = "";
int g_device = -1;

// Device memory of the allocation-free measurement (qcd_measure_noalloc):
// its partial sums and per-site buffers live in this file-scope array,
// which every GPU holds once, at its own address, until the process ends.
// Host code never uses the name `scratch` as an address: qcd_device_open
// looks up each device's copy with hipGetSymbolAddress and keeps it in
// g_scratch[device].
constexpr size_t kScratchBytes = kScratchBytesDev;
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(QCD_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

uint32_t align_up(uint32_t x, uint32_t a) { return (1u + (x - 1u) / a) * a; }

// Polyakov slice records: max_measurements x (N1 + N2 + N3) x (Re, Im)
size_t slice_rec_doubles(const qcd_lattice* L)
{
    return (size_t)2 * (L->p.n[0] + L->p.n[1] + L->p.n[2]) * (size_t)L->p.max_measurements;
}

// Optional per-launch extras: fused plaquette measurement and a pair of HIP
// events recorded at the start and end of the kernel itself
// (hipExtLaunchKernelGGL), used to time kernels without launch overhead.
struct LaunchOpt {
    int meas_mode = 0;
    hipEvent_t e0 = nullptr;
    hipEvent_t e1 = nullptr;
    uint32_t lds = 0;   // extra dynamic LDS per workgroup (caps workgroups per CU)
};

template <typename Kern, typename... Args>
void launch(Kern k, dim3 grid, dim3 block, hipStream_t s, const LaunchOpt& o, Args... args)
{
    if (o.e0) hipExtLaunchKernelGGL(k, grid, block, o.lds, s, o.e0, o.e1, 0, args...);
    else hipLaunchKernelGGL(k, grid, block, o.lds, s, args...);
}

// Times what is enqueued on the lattice stream during its lifetime under
// `name` when profiling is on (events before and after, then a wait).
struct ProfScope {
    qcd_lattice* L;
    const char* name;
    explicit ProfScope(qcd_lattice* l, const char* n) : L(l), name(n)
    {
        if (L->profiling) (void)hipEventRecord(L->prof_ev[0], L->stream);
    }
    ~ProfScope()
    {
        if (!L->profiling) return;
        float ms = 0.f;
        if (hipEventRecord(L->prof_ev[1], L->stream) != hipSuccess || hipEventSynchronize(L->prof_ev[1]) != hipSuccess ||
            hipEventElapsedTime(&ms, L->prof_ev[0], L->prof_ev[1]) != hipSuccess)
            return;
        for (auto& e : L->prof)
            if (e.name == name) {
                ++e.n;
                e.sum += ms;
                e.sumsq += (double)ms * ms;
                return;
            }
        L->prof.push_back({name, 1u, (double)ms, (double)ms * ms});
    }
};

// 64-link units one update launch covers, and the first of them: the whole
// lattice, or an x-slab's interior slices 1 .. N1 - 2 (x is the slowest
// coordinate of the compact index, so they are one contiguous range).
uint32_t update_units(const qcd_lattice* L)
{
    if (L->slab) return (L->g.vh / (uint32_t)L->g.n1) * (uint32_t)(L->g.n1 - 2) / kLinks;
    return align_up(L->g.vh, kLinks) / kLinks;
}
uint32_t update_unit0(const qcd_lattice* L) { return L->slab ? L->g.vh / (uint32_t)L->g.n1 / kLinks : 0u; }

// The unit -> XCD map the update kernels use when `want` is asked for;
// sets a.xcd_map to the mode that applies and the block parameters.
// Modes 2 and 3 need whole z-rows per 64-link unit (N2 / 2 divides 64),
// whole z-groups and whole units per XCD block; mode 2 takes x-slabs of
// N1 / 8 x-values, mode 3 the block (N1 / ax) x (N3 / az) z-rows,
// ax az = 8, with the least halo per link: an update reads the x and z
// neighbours of its links, so an XCD that owns a block fetches its halo
// slices once more than the compulsory traffic, about (1 / x-extent +
// 1 / z-extent) of the table for a block that is cut along both
// (periodic full extents have no halo).  A mode that does not apply falls
// back to mode 1 (and to 0 when shape1 needs whole workgroups of four
// units per XCD and does not get them).
void dispatch_order(const qcd_lattice* L, int want, bool shape1, UpdArgs& a)
{
    if (L->slab && want > 1) want = 1;   // the block orders cover the whole lattice
    a.xcd_map = want;
    a.ord_ng = a.ord_nt = a.ord_xs = a.ord_ax = a.ord_zr = 1;
    if (want >= 2) {
        const uint32_t h = (uint32_t)L->g.h, nb = L->g.vh / kLinks;
        const uint32_t rb = (h <= kLinks && kLinks % h == 0) ? kLinks / h : 0;
        bool ok = rb && (L->g.vh % kLinks == 0) && (nb % 8 == 0) && ((uint32_t)L->g.n3 % rb == 0);
        if (ok) {
            const uint32_t ng = (uint32_t)L->g.n3 / rb, n1 = (uint32_t)L->g.n1;
            double best = 1e30;
            uint32_t bax = 0;
            for (uint32_t ax = 8; ax >= 1; ax /= 2) {
                const uint32_t az = 8 / ax;
                if (n1 % ax || ng % az) continue;
                if (want == 2 && ax != 8) continue;
                const double cost = (ax > 1 ? (double)ax / n1 : 0.0) + (az > 1 ? (double)az / L->g.n3 : 0.0);
                if (cost < best - 1e-12) {
                    best = cost;
                    bax = ax;
                }
            }
            ok = bax != 0;
            if (ok) {
                a.ord_ng = ng;
                a.ord_nt = (uint32_t)L->g.n4;
                a.ord_ax = bax;
                a.ord_xs = n1 / bax;
                a.ord_zr = ng / (8 / bax);
            }
        }
        if (!ok) a.xcd_map = 1;
    }
    if (shape1 && a.xcd_map != 0) {
        // four 64-link units per workgroup: the XCD blocks need whole workgroups
        if (update_units(L) % (8u * kWaves1)) a.xcd_map = 0;
    }
}

template <typename R, int N, int NREAL, bool OVER, int RK>
void launch_update_t(qcd_lattice* L, const UpdArgs& a, const LaunchOpt& o)
{
    const uint32_t nb = update_units(L);   // 64-link units
    if (a.tw.on) {   // twisted boundary: three-wavefront shape, plaquette never fused
        launch(k_update<R, N, NREAL, OVER, RK, false, true>, dim3(nb), dim3(kUpdThreads), L->stream, o, a);
        return;
    }
    if (L->tune_update_shape == 1) {
        const uint32_t grid = align_up(nb, kWaves1) / kWaves1;
        if (o.meas_mode)
            launch(k_update1<R, N, NREAL, OVER, RK, true>, dim3(grid), dim3(kWaves1 * kLinks), L->stream, o, a, nb);
        else
            launch(k_update1<R, N, NREAL, OVER, RK, false>, dim3(grid), dim3(kWaves1 * kLinks), L->stream, o, a, nb);
        return;
    }
    if (o.meas_mode) {
        launch(k_update<R, N, NREAL, OVER, RK, true>, dim3(nb), dim3(kUpdThreads), L->stream, o, a);
        return;
    }
    // the plain half-steps, most launches of a cycle: one instance per direction
    const dim3 grid(nb), block(kUpdThreads);
    switch (a.mu) {
        case 0: launch(k_update<R, N, NREAL, OVER, RK, false, false, 0>, grid, block, L->stream, o, a); break;
        case 1: launch(k_update<R, N, NREAL, OVER, RK, false, false, 1>, grid, block, L->stream, o, a); break;
        case 2: launch(k_update<R, N, NREAL, OVER, RK, false, false, 2>, grid, block, L->stream, o, a); break;
        default: launch(k_update<R, N, NREAL, OVER, RK, false, false, 3>, grid, block, L->stream, o, a); break;
    }
}

// One produce() of `samples` quads per generator instance into `out` (quad
// q = sample * n_inst + instance), or a skip of as many when out is null;
// XOR128 / PM states advance in place in the current state buffer.
void produce_quads_into(qcd_lattice* L, uint32_t samples, float4* out)
{
    const uint32_t n = L->n_inst, nb = align_up(n, kBlock) / kBlock;
    if (L->p.prng == QCD_PRNG_RANLUX) {
        launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, n, samples, out, L->stream);
    } else if (L->p.prng == QCD_PRNG_XOR128) {
        if (out)
            hipLaunchKernelGGL(k_xor128_produce, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n,
                               samples, out);
        else
            hipLaunchKernelGGL(k_xor128_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, samples);
    } else {
        if (out)
            hipLaunchKernelGGL(k_pm_produce, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, samples,
                               out);
        else
            hipLaunchKernelGGL(k_pm_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, samples);
    }
}

template <typename R, int N, int NREAL>
int launch_update_rn(qcd_lattice* L, int parity, int mu, int kind, int start_gid, const LaunchOpt& o)
{
    UpdArgs a;
    a.table = L->table;
    a.src = L->table;
    if (L->shadow) {   // race check: the gathers read a copy taken before the half-step
        HIPCHK(hipMemcpyAsync(L->shadow, L->table, L->table_bytes, hipMemcpyDeviceToDevice, L->stream));
        a.src = L->shadow;
    }
    a.g = L->g;
    a.parity = parity;
    a.mu = mu;
    a.beta = L->p.beta / L->p.sun;
    a.nhit = L->p.nhit;
    a.sites = 2u * L->g.vh;
    a.rng = L->rng_buf[L->rng_cur];
    a.rng_out = L->rng_buf[L->rng_cur ^ 1];
    a.jump = L->jump_cols;
    a.prn = nullptr;
    a.prn_stride = 0;
    a.pm_ak = L->pm_ak;
    a.s_quads = (L->p.stream_mode == QCD_STREAM_REFERENCE) ? L->s_quads : 0u;
    a.partials = L->part_fused;
    a.meas_mode = o.meas_mode;
    a.unit0 = update_unit0(L);
    a.tw = L->tw;
    dispatch_order(L, L->tune_xcd_map, L->tune_update_shape == 1 && !a.tw.on, a);
    LaunchOpt ol = o;
    if (L->tune_blocks_per_cu > 0) {
        // 160 KiB of LDS per CU on gfx950; reserve so at most k workgroups fit.
        const uint32_t per = (160u * 1024u) / (uint32_t)L->tune_blocks_per_cu;
        const uint32_t stat = (uint32_t)(2 * 2 * N * N * kLinks * sizeof(R));
        ol.lds = per > stat + 256 ? per - stat - 256 : 0;
    }
    if (kind == QCD_OVERRELAX) {
        ProfScope ps(L, o.meas_mode ? "overrelaxation half-step + plaquette" : "overrelaxation half-step");
        launch_update_t<R, N, NREAL, true, RK_NONE>(L, a, ol);
    } else if (start_gid) {
        ProfScope ps(L, "heatbath half-step (GID numbers)");
        launch_update_t<R, N, NREAL, false, RK_GID>(L, a, ol);
    } else {
        if (!L->rng_ready) return fail(QCD_E_STATE, "heatbath needs qcd_prng_init first");
        if (L->quad_buffer) {
            // the reference's produce() into the PRN buffer, then the update
            // reads quad k of link i at k * PRNGSTEP + i
            {
                ProfScope ps(L, "PRNG produce");
                produce_quads_into(L, L->s_quads, L->prn_buf);
            }
            a.prn = L->prn_buf;
            a.prn_stride = L->n_streams;
            ProfScope ps(L, o.meas_mode ? "heatbath half-step + plaquette" : "heatbath half-step");
            launch_update_t<R, N, NREAL, false, RK_BUF>(L, a, ol);
        } else {
            ProfScope ps(L, o.meas_mode ? "heatbath half-step + plaquette" : "heatbath half-step");
            if (L->p.prng == QCD_PRNG_XOR128) launch_update_t<R, N, NREAL, false, RK_XOR128>(L, a, ol);
            else if (L->p.prng == QCD_PRNG_PM) launch_update_t<R, N, NREAL, false, RK_PM>(L, a, ol);
            else return fail(QCD_E_UNSUPPORTED, "generator %d not available for fused updates", L->p.prng);
            L->rng_cur ^= 1;   // the kernel wrote the states into the other buffer
        }
        ++L->prng_counter;
    }
    HIPCHK(hipGetLastError());
    return QCD_OK;
}

#define DISPATCH(L, FN, ...)                                                                          \
    ((L)->p.precision == QCD_FP32                                                                     \
         ? ((L)->p.sun == 3 ? ((L)->p.storage_reals == 12 ? FN<float, 3, 12>(__VA_ARGS__)            \
                                                          : FN<float, 3, 18>(__VA_ARGS__))           \
                            : FN<float, 2, 4>(__VA_ARGS__))                                           \
         : ((L)->p.sun == 3 ? ((L)->p.storage_reals == 12 ? FN<double, 3, 12>(__VA_ARGS__)           \
                                                          : FN<double, 3, 18>(__VA_ARGS__))          \
                            : FN<double, 2, 4>(__VA_ARGS__)))

template <typename R, int N, int NREAL>
int launch_simple(qcd_lattice* L, int which, const LaunchOpt& o)
{
    uint32_t nb = align_up(L->g.vh, kBlock) / kBlock;
    switch (which) {
        case 0:
            hipLaunchKernelGGL((k_init_cold<R, N, NREAL>), dim3(nb), dim3(kBlock), 0, L->stream, L->table, L->g);
            break;
        case 1:
            hipLaunchKernelGGL((k_init_gid<R, N, NREAL>), dim3(nb), dim3(kBlock), 0, L->stream, L->table, L->g,
                               2u * L->g.vh);
            break;
        case 2:
            hipLaunchKernelGGL((k_reunitarize<R, N, NREAL>), dim3(nb), dim3(kBlock), 0, L->stream, L->table, L->g);
            break;
        case 3:
        {
            double* sp = L->want_aslices ? L->site_pl : nullptr;
            if (L->tw.on)
                launch(k_plaquette<R, N, NREAL, true>, dim3(L->upd_blocks, 2), dim3(kBlock), L->stream, o, L->table,
                       L->g, L->part_plq, L->tw, sp);
            else
                launch(k_plaquette<R, N, NREAL>, dim3(L->upd_blocks, 2), dim3(kBlock), L->stream, o, L->table, L->g,
                       L->part_plq, L->tw, sp);
            if (sp)
                hipLaunchKernelGGL(k_slices, dim3((unsigned)(L->g.n1 + L->g.n2 + L->g.n3)), dim3(kBlock), 0,
                                   L->stream, (const double*)sp, L->g, (uint32_t)L->g.n4, L->a_rec,
                                   L->p.max_measurements, (const int*)L->rec_index, 0);
        }
            break;
        case 4:
            launch(k_polyakov<R, N, NREAL>, dim3(L->poly_blocks), dim3(kBlock), L->stream, o, L->table, L->g,
                   L->part_poly, L->want_slices ? L->site_tr : (double*)nullptr);
            if (L->want_slices)
                hipLaunchKernelGGL(k_slices, dim3((unsigned)(L->g.n1 + L->g.n2 + L->g.n3)), dim3(kBlock), 0,
                                   L->stream, (const double*)L->site_tr, L->g, 1u, L->slice_rec,
                                   L->p.max_measurements, (const int*)L->rec_index, 3);
            break;
        case 6:
            if (L->tw.on)
                hipLaunchKernelGGL((k_wilson<R, N, NREAL, true>), dim3(L->wil_blocks), dim3(kBlock), 0, L->stream,
                                   L->table, L->g, L->wil_R, L->wil_T, L->part_wil, L->tw);
            else
                hipLaunchKernelGGL((k_wilson<R, N, NREAL>), dim3(L->wil_blocks), dim3(kBlock), 0, L->stream, L->table,
                                   L->g, L->wil_R, L->wil_T, L->part_wil, L->tw);
            break;
        case 7:
            if (L->tw.on)
                hipLaunchKernelGGL((k_fmunu<R, N, NREAL, true>), dim3(L->upd_blocks, 2), dim3(kBlock), 0, L->stream,
                                   L->table, L->g, L->part_f, L->f_efield, L->f_gen[0], L->f_gen[1], L->tw);
            else
                hipLaunchKernelGGL((k_fmunu<R, N, NREAL>), dim3(L->upd_blocks, 2), dim3(kBlock), 0, L->stream,
                                   L->table, L->g, L->part_f, L->f_efield, L->f_gen[0], L->f_gen[1], L->tw);
            break;
        case 5:
            if (L->p.prng == QCD_PRNG_XOR128) {
                hipLaunchKernelGGL((k_init_hot<R, N, NREAL, RK_XOR128>), dim3(nb), dim3(kBlock), 0, L->stream, L->table,
                                   L->g, L->rng_buf[L->rng_cur], (const float4*)nullptr, 0u);
            } else if (L->p.prng == QCD_PRNG_PM) {
                hipLaunchKernelGGL((k_init_hot<R, N, NREAL, RK_PM>), dim3(nb), dim3(kBlock), 0, L->stream, L->table,
                                   L->g, L->rng_buf[L->rng_cur], (const float4*)nullptr, 0u);
            } else if (L->p.prng == QCD_PRNG_RANLUX) {
                const uint32_t q = (N == 3) ? 24u : 8u;   // 8 links x 3 or 1 quads
                launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, L->n_streams, q, L->prn_buf, L->stream);
                hipLaunchKernelGGL((k_init_hot<R, N, NREAL, RK_BUF>), dim3(nb), dim3(kBlock), 0, L->stream, L->table,
                                   L->g, (uint4*)nullptr, (const float4*)L->prn_buf, L->n_streams);
            } else {
                return fail(QCD_E_UNSUPPORTED, "generator %d not available for the hot start", L->p.prng);
            }
            break;
        default:
            return fail(QCD_E_ARG, "internal: bad launch %d", which);
    }
    HIPCHK(hipGetLastError());
    return QCD_OK;
}

template <typename R, int N, int NREAL>
int launch_random(qcd_lattice* L, const float4* buf, uint32_t rowsize, int d)
{
    const uint32_t nb = align_up(L->g.vh, kBlock) / kBlock;
    hipLaunchKernelGGL((k_init_random<R, N, NREAL>), dim3(nb), dim3(kBlock), 0, L->stream, L->table, L->g, buf,
                       rowsize, d);
    HIPCHK(hipGetLastError());
    return QCD_OK;
}

// flags: 1 = fused plaquette partials, 2 = k_plaquette partials, 4 = Polyakov.
FinArgs fin_args(const qcd_lattice* L, int flags)
{
    FinArgs f;
    f.fused = (flags & 1) ? L->part_fused : nullptr;
    f.n_fused = L->nb_fused;
    f.plq = (flags & 2) ? L->part_plq : nullptr;
    f.n_plq = 2u * L->upd_blocks;
    f.poly = (flags & 4) ? L->part_poly : nullptr;
    f.n_poly = L->poly_blocks;
    f.wil = (flags & 8) ? L->part_wil : nullptr;
    f.n_wil = L->wil_blocks;
    f.plq_rec = L->plq_rec;
    f.poly_rec = L->poly_rec;
    f.wil_rec = L->wil_rec;
    f.rec_index = L->rec_index;
    f.capacity = L->p.max_measurements;
    f.slices = (flags & 16) ? 1 : 0;
    f.fmunu = (flags & 32) ? L->part_f : nullptr;
    f.f_rec = L->f_rec;
    f.a_rec = L->a_rec;
    f.a_n = L->g.n1 + L->g.n2 + L->g.n3;
    f.aslices = (flags & 64) ? 1 : 0;
    return f;
}

int launch_finalize(qcd_lattice* L, int flags, const LaunchOpt& o)
{
    launch(k_finalize, dim3(1), dim3(kBlock), L->stream, o, fin_args(L, flags));
    HIPCHK(hipGetLastError());
    return QCD_OK;
}

int ensure_device()
{
    if (g_device < 0) {
        int rc = qcd_device_open(0, 0);
        if (rc) return rc;
    }
    return QCD_OK;
}

}  // namespace

int qcdhip::set_error(int code, const char* msg)
{
    return fail(code, "%s", msg);
}

int qcdhip::write_fmunu(qcd_lattice* L, const double* f, int count)
{
    if (!L || count < 0 || (count && !f)) return fail(QCD_E_ARG, "bad F_munu records");
    const int m = L->p.max_measurements;
    const int n = count < m ? count : m;
    HIPCHK(hipStreamSynchronize(L->stream));
    if (n) HIPCHK(hipMemcpy(L->f_rec, f, sizeof(double) * 12 * n, hipMemcpyHostToDevice));
    return QCD_OK;
}

int qcdhip::su2_quats(qcd_lattice* L, float4* q, bool to_quats, hipStream_t* stream)
{
    if (!L || !q) return fail(QCD_E_ARG, "null argument");
    if (L->p.sun != 2 || L->p.precision != QCD_FP32)
        return fail(QCD_E_UNSUPPORTED, "quaternion links: SU(2) lattices in FP32 only");
    const dim3 grid(align_up(L->g.vh, kBlock) / kBlock, 8);
    if (to_quats) hipLaunchKernelGGL(k_su2_to_quats, grid, dim3(kBlock), 0, L->stream, L->table, L->g, q);
    else hipLaunchKernelGGL(k_quats_to_su2, grid, dim3(kBlock), 0, L->stream, L->table, L->g, (const float4*)q);
    HIPCHK(hipGetLastError());
    if (stream) *stream = L->stream;
    return QCD_OK;
}

int qcdhip::lattice_dims(const qcd_lattice* L, int n[4])
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    for (int k = 0; k < 4; ++k) n[k] = L->p.n[k];
    return QCD_OK;
}

// ---------------------------------------------------------------- C ABI
extern "C" {

const char* qcd_last_error(void) { return g_err; }
int qcd_abi_version(void) { return QCD_ABI_VERSION; }

int qcd_device_open(int platform, int device)
{
    if (platform != 0) return fail(QCD_E_ARG, "platform %d: only platform 0 (HIP) exists", platform);
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(QCD_E_ARG, "device %d not present (%d devices)", device, count);
    HIPCHK(hipSetDevice(device));
    if (device < kMaxDevices && !g_scratch[device]) {
        // this device's copy of the allocation-free measurement's array
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(QCD_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[device] = (double*)ptr;
    }
    g_device = device;
    return QCD_OK;
}

int qcd_device_close(void)
{
    if (g_device >= 0) HIPCHK(hipDeviceSynchronize());
    g_device = -1;
    return QCD_OK;
}

int qcd_device_info(char* name, int name_len, int* compute_units, size_t* global_mem)
{
    int rc = ensure_device();
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, g_device));
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (global_mem) *global_mem = prop.totalGlobalMem;
    return QCD_OK;
}

// Quads per instance of one produce() (suncl.cpp:2302-2307,
// random.cpp:219-223): align64(3 PRNGSTEP (NHIT+1) NHITPar) quads in all,
// or, in a run that starts hot (model_start_hot), align64(ROWSIZE ND +
// 3 PRNGSTEP (NHIT+1)) with ROWSIZE = align64(V), over PRNG_instances
// (PRNGSTEP = align64(V/2) unless the parameters set another count),
// rounded up: PRNG_samples = 1 + (total - 1) / PRNG_instances.
static uint32_t produce_quads(const qcd_params& p, uint32_t vh, bool hot_run)
{
    auto align64 = [](uint64_t x) { return (x + 63u) / 64u * 64u; };
    const uint64_t prngstep = align64(vh);
    const uint64_t inst = p.prng_instances ? (uint64_t)p.prng_instances : prngstep;
    const uint64_t rowsize = align64(2ull * vh);
    const uint64_t upd = 3ull * prngstep * (uint64_t)(p.nhit + 1);
    const uint64_t total = align64(hot_run ? rowsize * 4u + upd : (uint64_t)p.nhitpar * upd);
    return (uint32_t)(1 + (total - 1) / inst);
}

// Reference-stream skip of s_quads quads per produce(): XOR128 jump matrix
// for 4 * s_quads steps (four steps per quad), Park-Miller multiplier for
// s_quads steps (pm_quad steps each of its four lanes once).  Columns (image
// of state bit b = word b/32, bit b%32), then the four-bit table: entry
// [c][e] = XOR of the columns of the bits of chunk c (bits 4c .. 4c+3) set
// in e.
static int build_jumps(qcd_lattice* L)
{
    uint4 cols[kJumpCols + kJumpTab];
    for (int b = 0; b < kJumpCols; ++b) {
        uint4 st = make_uint4(0u, 0u, 0u, 0u);
        uint32_t* w = (uint32_t*)&st;
        w[b / 32] = 1u << (b % 32);
        for (uint32_t k = 0; k < 4u * L->s_quads; ++k) (void)xor128_step(st);
        cols[b] = st;
    }
    for (int c = 0; c < 32; ++c)
        for (int e = 0; e < 16; ++e) {
            uint4 acc = make_uint4(0u, 0u, 0u, 0u);
            for (int k = 0; k < 4; ++k)
                if ((e >> k) & 1) {
                    const uint4 v = cols[4 * c + k];
                    acc.x ^= v.x;
                    acc.y ^= v.y;
                    acc.z ^= v.z;
                    acc.w ^= v.w;
                }
            cols[kJumpCols + 16 * c + e] = acc;
        }
    if (hipMemcpy(L->jump_cols, cols, sizeof(cols), hipMemcpyHostToDevice) != hipSuccess)
        return fail(QCD_E_HIP, "jump matrix upload failed");
    uint32_t ak = 1;
    for (uint32_t k = 0; k < L->s_quads; ++k) ak = pm_mulmod(ak, 16807u);
    L->pm_ak = ak;
    return QCD_OK;
}

// Switches the produce() size between the cold/GID runs' and the hot runs'
// (QCD_START_RANDOM).  A change rebuilds the jump tables, grows the RANLUX
// quad buffer when needed and drops a captured cycle (its launches hold the
// old size); the caller re-seeds streams seeded under the old size.
static int set_sizing(qcd_lattice* L, bool hot_run, bool* changed)
{
    *changed = false;
    const uint32_t S = produce_quads(L->p, L->g.vh, hot_run);
    if (S == L->s_quads && hot_run == L->hot_run) return QCD_OK;
    if (L->batch_owned)
        return fail(QCD_E_STATE, "a batched chain keeps its stream sizing (its batch holds the buffers): "
                                 "no reference hot start or resume of a hot run on it");
    HIPCHK(hipStreamSynchronize(L->stream));
    *changed = true;
    L->hot_run = hot_run;
    L->s_quads = S;
    drop_cycle_graphs(L);
    if (L->quad_buffer && L->prn_quads < S) {
        (void)hipFree(L->prn_buf);
        L->prn_buf = nullptr;
        L->prn_quads = S;
        if (hipMalloc((void**)&L->prn_buf, sizeof(float4) * (size_t)S * L->n_inst) != hipSuccess)
            return fail(QCD_E_NOMEM, "PRN quad buffer of %u quads per instance", S);
    }
    return build_jumps(L);
}

int qcd_lattice_create(const qcd_params* pp, qcd_lattice** out)
{
    if (!pp || !out) return fail(QCD_E_ARG, "null argument");
    const qcd_params& p = *pp;
    if (p.sun != 2 && p.sun != 3) return fail(QCD_E_ARG, "sun must be 2 or 3");
    for (int d = 0; d < 4; ++d)
        if (p.n[d] < 2 || (p.n[d] & 1)) return fail(QCD_E_ARG, "extent N%d = %d must be even and >= 2", d + 1, p.n[d]);
    if (p.precision != QCD_FP32 && p.precision != QCD_FP64) return fail(QCD_E_ARG, "bad precision");
    if (p.sun == 3 && p.storage_reals != 18 && p.storage_reals != 12)
        return fail(QCD_E_ARG, "SU(3) storage must be 18 or 12 reals");
    if (p.sun == 2 && p.storage_reals != 4) return fail(QCD_E_ARG, "SU(2) storage must be 4 reals");
    if (p.nhit < 1) return fail(QCD_E_ARG, "nhit must be >= 1");
    if (p.nhitpar != 1) return fail(QCD_E_UNSUPPORTED, "NHITPar other than 1 is not implemented");
    if (p.max_measurements < 1) return fail(QCD_E_ARG, "max_measurements must be >= 1");
    if (p.prng_instances % 64u)
        return fail(QCD_E_ARG, "prng_instances = %u: 0 or a multiple of 64", p.prng_instances);
    if (p.prng_instances && p.stream_mode != QCD_STREAM_REFERENCE)
        return fail(QCD_E_ARG, "prng_instances needs the reference stream layout");
    uint64_t V = (uint64_t)p.n[0] * p.n[1] * p.n[2] * p.n[3];
    if (V / 2 > 0x7fffffffull) return fail(QCD_E_ARG, "lattice too large");
    int rc = ensure_device();
    if (rc) return rc;

    qcd_lattice* L = new qcd_lattice();
    L->p = p;
    if (p.sun == 2) L->f_gen[1] = 2;   // SU(2): sigma_3 and sigma_2
    L->g.n1 = p.n[0];
    L->g.n2 = p.n[1];
    L->g.n3 = p.n[2];
    L->g.n4 = p.n[3];
    L->g.h = p.n[1] / 2;
    L->g.dh = make_fastdiv((uint32_t)L->g.h);
    L->g.d3 = make_fastdiv((uint32_t)p.n[2]);
    L->g.d4 = make_fastdiv((uint32_t)p.n[3]);
    L->g.vh = (uint32_t)(V / 2);
    // Plane stride: a multiple of 64 sites (512 B / 1 KiB, whole cache lines
    // per wavefront) plus one extra 64-site unit, so the 72 planes of one
    // link do not share an L2 set/channel when V/2 is a power of two (for
    // 32^3 x 8 a bare stride would be exactly 1 MiB).
    L->g.vh_pad = align_up(L->g.vh, 64) + 64;
    L->np = p.storage_reals / 2;
    L->elem = (p.precision == QCD_FP32) ? 8 : 16;
    L->table_bytes = (size_t)2 * 4 * L->np * L->g.vh_pad * L->elem;
    L->upd_blocks = align_up(L->g.vh, kBlock) / kBlock;
    uint32_t vs = (uint32_t)p.n[0] * p.n[1] * p.n[2];
    L->poly_blocks = align_up(8u * vs, kBlock) / kBlock;   // 8 lanes per spatial site
    L->n_streams = align_up(L->g.vh, 64);
    L->n_inst = p.prng_instances ? p.prng_instances : L->n_streams;
    L->quad_buffer = (p.prng == QCD_PRNG_RANLUX) || p.prng_instances != 0;
    L->s_quads = produce_quads(p, L->g.vh, false);

    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete L;
        return fail(QCD_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    L->nb_fused = align_up(L->g.vh, kLinks) / kLinks;
    L->wil_blocks = align_up(2u * L->g.vh, kBlock) / kBlock;
    size_t part_doubles = (size_t)3 * L->nb_fused + (size_t)2 * 2 * L->upd_blocks + (size_t)4 * L->poly_blocks +
                          (size_t)L->wil_blocks;
    if (hipMalloc(&L->table, L->table_bytes) != hipSuccess ||
        hipMalloc((void**)&L->rng_buf[0], sizeof(uint4) * L->n_inst) != hipSuccess ||
        hipMalloc((void**)&L->rng_buf[1], sizeof(uint4) * L->n_inst) != hipSuccess ||
        hipMalloc((void**)&L->jump_cols, sizeof(uint4) * (kJumpCols + kJumpTab)) != hipSuccess ||
        hipMalloc((void**)&L->partials, sizeof(double) * part_doubles) != hipSuccess ||
        hipMalloc((void**)&L->rec_index, sizeof(int) * 8) != hipSuccess ||
        hipMalloc((void**)&L->plq_rec, sizeof(double) * 2 * p.max_measurements) != hipSuccess ||
        hipMalloc((void**)&L->poly_rec, sizeof(double) * 4 * p.max_measurements) != hipSuccess ||
        hipMalloc((void**)&L->wil_rec, sizeof(double) * p.max_measurements) != hipSuccess ||
        hipMalloc((void**)&L->site_tr, sizeof(double) * 2 * vs) != hipSuccess ||
        hipMalloc((void**)&L->slice_rec, sizeof(double) * slice_rec_doubles(L)) != hipSuccess ||
        hipMalloc((void**)&L->a_rec, sizeof(double) * slice_rec_doubles(L)) != hipSuccess ||
        hipMalloc((void**)&L->part_f, sizeof(double) * 12 * 2 * L->upd_blocks) != hipSuccess ||
        hipMalloc((void**)&L->f_rec, sizeof(double) * 12 * p.max_measurements) != hipSuccess) {
        qcd_lattice_destroy(L);
        return fail(QCD_E_NOMEM, "device allocation failed (table %zu bytes)", L->table_bytes);
    }
    if (L->quad_buffer) {
        // One produce() per heatbath half-step (s_quads per instance) or, with
        // PRNGSTEP instances, per Haar start (8 links x 3 quads for SU(3)).
        L->prn_quads = L->s_quads > 24u ? L->s_quads : 24u;
        if ((p.prng == QCD_PRNG_RANLUX &&
             hipMalloc((void**)&L->rl_state, sizeof(float4) * 7 * L->n_inst) != hipSuccess) ||
            hipMalloc((void**)&L->prn_buf, sizeof(float4) * (size_t)L->prn_quads * L->n_inst) != hipSuccess) {
            qcd_lattice_destroy(L);
            return fail(QCD_E_NOMEM, "PRN quad buffer");
        }
    }
    L->part_fused = L->partials;
    L->part_plq = L->part_fused + (size_t)3 * L->nb_fused;
    L->part_poly = L->part_plq + (size_t)2 * 2 * L->upd_blocks;
    L->part_wil = L->part_poly + (size_t)4 * L->poly_blocks;
    (void)hipMemset(L->rec_index, 0, sizeof(int) * 8);
    (void)hipMemset(L->table, 0, L->table_bytes);
    {   // records start at zero, so state images of equal runs are equal byte for byte
        const size_t m = (size_t)p.max_measurements;
        (void)hipMemset(L->plq_rec, 0, sizeof(double) * 2 * m);
        (void)hipMemset(L->poly_rec, 0, sizeof(double) * 4 * m);
        (void)hipMemset(L->wil_rec, 0, sizeof(double) * m);
        (void)hipMemset(L->slice_rec, 0, sizeof(double) * slice_rec_doubles(L));
        (void)hipMemset(L->a_rec, 0, sizeof(double) * slice_rec_doubles(L));
        (void)hipMemset(L->f_rec, 0, sizeof(double) * 12 * m);
    }
    if ((rc = build_jumps(L)) != QCD_OK) {
        qcd_lattice_destroy(L);
        return rc;
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        qcd_lattice_destroy(L);
        return fail(QCD_E_HIP, "initial clear failed");
    }
    *out = L;
    return QCD_OK;
}

int qcd_lattice_destroy(qcd_lattice* L)
{
    if (!L) return QCD_OK;
    if (L->batch_owned) return fail(QCD_E_STATE, "a batched chain is destroyed with its batch (qcd_batch_destroy)");
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    drop_cycle_graphs(L);
    if (L->side) (void)hipStreamDestroy(L->side);
    if (L->ev_fork) (void)hipEventDestroy(L->ev_fork);
    if (L->ev_join) (void)hipEventDestroy(L->ev_join);
    for (auto& e : L->prof_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(L->table);
    (void)hipFree(L->shadow);
    (void)hipFree(L->rng_buf[0]);
    (void)hipFree(L->rng_buf[1]);
    (void)hipFree(L->jump_cols);
    (void)hipFree(L->rl_state);
    (void)hipFree(L->prn_buf);
    (void)hipFree(L->partials);
    (void)hipFree(L->rec_index);
    (void)hipFree(L->plq_rec);
    (void)hipFree(L->poly_rec);
    (void)hipFree(L->wil_rec);
    (void)hipFree(L->site_tr);
    (void)hipFree(L->slice_rec);
    (void)hipFree(L->a_rec);
    (void)hipFree(L->site_pl);
    (void)hipFree(L->part_f);
    (void)hipFree(L->f_rec);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
    return QCD_OK;
}

int qcd_lattice_sync(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    HIPCHK(hipStreamSynchronize(L->stream));
    return QCD_OK;
}

// Kernel names of the reference's lattice_make_programs (suncl.cpp:3265-3648)
// and PRNG::initialize (random.cpp:487-493,668,886), each bound to what this
// library runs for it.  Names whose work is done inside another launch here
// (the PRN production inside every heatbath update, the reductions inside
// the measurement's final sums, the four directions of the reference's hot
// start in one call) are accepted and run nothing.
enum KAct { KA_NONE, KA_HB, KA_OR, KA_COLD, KA_GID, KA_HOT, KA_GS, KA_MEAS };
struct KName {
    const char* name;
    KAct act;
    int parity, mu, mask;
};
static const KName kKernelNames[] = {
    {"update_odd_X", KA_HB, 1, 0, 0},   {"update_odd_Y", KA_HB, 1, 1, 0},   {"update_odd_Z", KA_HB, 1, 2, 0},
    {"update_odd_T", KA_HB, 1, 3, 0},   {"update_even_X", KA_HB, 0, 0, 0},  {"update_even_Y", KA_HB, 0, 1, 0},
    {"update_even_Z", KA_HB, 0, 2, 0},  {"update_even_T", KA_HB, 0, 3, 0},
    // overrelaxation half-steps: no reference kernel, named after the updates
    {"overrelax_odd_X", KA_OR, 1, 0, 0}, {"overrelax_odd_Y", KA_OR, 1, 1, 0}, {"overrelax_odd_Z", KA_OR, 1, 2, 0},
    {"overrelax_odd_T", KA_OR, 1, 3, 0}, {"overrelax_even_X", KA_OR, 0, 0, 0}, {"overrelax_even_Y", KA_OR, 0, 1, 0},
    {"overrelax_even_Z", KA_OR, 0, 2, 0}, {"overrelax_even_T", KA_OR, 0, 3, 0},
    {"lattice_init_cold", KA_COLD, 0, 0, 0}, {"lattice_init_gid", KA_GID, 0, 0, 0},
    {"lattice_init_hot_X", KA_HOT, 0, 0, 0}, {"lattice_init_hot_Y", KA_NONE, 0, 0, 0},
    {"lattice_init_hot_Z", KA_NONE, 0, 0, 0}, {"lattice_init_hot_T", KA_NONE, 0, 0, 0},
    {"lattice_GramSchmidt", KA_GS, 0, 0, 0},
    {"lattice_measurement_plq", KA_MEAS, 0, 0, QCD_MEAS_PLAQUETTE},
    {"lattice_measurement", KA_MEAS, 0, 0, QCD_MEAS_PLAQUETTE | QCD_MEAS_FMUNU},
    {"lattice_polyakov", KA_MEAS, 0, 0, QCD_MEAS_POLYAKOV},
    {"lattice_measurement_wilson", KA_MEAS, 0, 0, QCD_MEAS_WILSON},
    {"lattice_polyakov_diff_x", KA_MEAS, 0, 0, QCD_MEAS_POLYAKOV_SLICES},
    {"lattice_polyakov_diff_y", KA_NONE, 0, 0, 0}, {"lattice_polyakov_diff_z", KA_NONE, 0, 0, 0},
    {"lattice_action_diff_x", KA_MEAS, 0, 0, QCD_MEAS_ACTION_SLICES},
    {"lattice_action_diff_y", KA_NONE, 0, 0, 0}, {"lattice_action_diff_z", KA_NONE, 0, 0, 0},
    {"reduce_measurement_plq_double2", KA_NONE, 0, 0, 0}, {"reduce_measurement_double2", KA_NONE, 0, 0, 0},
    {"reduce_polyakov_double2", KA_NONE, 0, 0, 0}, {"reduce_wilson_double2", KA_NONE, 0, 0, 0},
    {"reduce_polyakov_diff_x_double2", KA_NONE, 0, 0, 0}, {"reduce_polyakov_diff_y_double2", KA_NONE, 0, 0, 0},
    {"reduce_polyakov_diff_z_double2", KA_NONE, 0, 0, 0}, {"reduce_action_diff_x_double2", KA_NONE, 0, 0, 0},
    {"reduce_action_diff_y_double2", KA_NONE, 0, 0, 0}, {"reduce_action_diff_z_double2", KA_NONE, 0, 0, 0},
    {"clear_measurement", KA_NONE, 0, 0, 0},
    {"xor128", KA_NONE, 0, 0, 0}, {"pm", KA_NONE, 0, 0, 0}, {"ranlux", KA_NONE, 0, 0, 0},
};

int qcd_kernel_create(qcd_lattice* L, const char* name, int* id)
{
    if (!L || !name || !id) return fail(QCD_E_ARG, "null argument");
    for (size_t k = 0; k < sizeof(kKernelNames) / sizeof(kKernelNames[0]); ++k)
        if (strcmp(kKernelNames[k].name, name) == 0) {
            L->kernels.push_back({(int)k, 0u});
            *id = (int)L->kernels.size();   // ids count from 1, as GPU::kernel_init's
            return QCD_OK;
        }
    return fail(QCD_E_ARG, "no kernel \"%s\" on this path", name);
}

int qcd_kernel_enqueue(qcd_lattice* L, int id)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (id < 1 || id > (int)L->kernels.size()) return fail(QCD_E_ARG, "kernel id %d was not created", id);
    const KName& k = kKernelNames[L->kernels[id - 1].first];
    int rc = QCD_OK;
    switch (k.act) {
        case KA_HB: rc = qcd_halfstep(L, k.parity, k.mu, QCD_HEATBATH); break;
        case KA_OR: rc = qcd_halfstep(L, k.parity, k.mu, QCD_OVERRELAX); break;
        case KA_COLD: rc = qcd_lattice_init(L, QCD_START_COLD); break;
        case KA_GID: rc = qcd_lattice_init(L, QCD_START_GID); break;
        case KA_HOT: rc = qcd_lattice_init(L, QCD_START_RANDOM); break;
        case KA_GS: rc = qcd_reunitarize(L); break;
        case KA_MEAS: rc = qcd_measure(L, k.mask); break;
        default: break;
    }
    if (rc == QCD_OK) ++L->kernels[id - 1].second;
    return rc;
}

int qcd_kernel_run(qcd_lattice* L, int id)
{
    const int rc = qcd_kernel_enqueue(L, id);
    if (rc) return rc;
    if (L->graph_open) return QCD_OK;   // under capture nothing runs yet
    HIPCHK(hipStreamSynchronize(L->stream));
    return QCD_OK;
}

int qcd_kernel_info(const qcd_lattice* L, int id, char* name, int name_len, unsigned* starts)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (id < 1 || id > (int)L->kernels.size()) return fail(QCD_E_ARG, "kernel id %d was not created", id);
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", kKernelNames[L->kernels[id - 1].first].name);
    if (starts) *starts = L->kernels[id - 1].second;
    return QCD_OK;
}

// Device buffers by integer id, as GPU::buffer_init / buffer_write /
// buffer_map / buffer_kill keep them (clinterface.cpp:1296-1440,1558-1568):
// a host written against that registry keeps its auxiliary buffers here.
// Types and their numbers follow GPU_buffer_types (clinterface.h:73-82).
}  // extern "C"

namespace {
struct DevBuffer {
    int type = 0;
    size_t bytes = 0;
    void* dev = nullptr;
    void* host = nullptr;               // the caller's host copy (buffer_write uploads it)
    std::vector<unsigned char> mirror;  // what qcd_buffer_map returns
    bool live = false;
};
std::vector<DevBuffer>& buffers()
{
    static std::vector<DevBuffer> b(1);   // ids count from 1
    return b;
}
DevBuffer* buffer_of(int id)
{
    auto& b = buffers();
    if (id <= 0 || id >= (int)b.size() || !b[id].live) return nullptr;
    return &b[id];
}
}  // namespace

extern "C" {

int qcd_buffer_create(int type, size_t count, size_t elem_size, void* host_ptr, int* id)
{
    if (!id) return fail(QCD_E_ARG, "null id");
    if (type < QCD_BUFFER_GLOBAL || type > QCD_BUFFER_UAV) return fail(QCD_E_ARG, "buffer type %d", type);
    if (count && elem_size > SIZE_MAX / count) return fail(QCD_E_ARG, "buffer size overflows");
    const bool copy = type == QCD_BUFFER_INPUT || type == QCD_BUFFER_IO || type == QCD_BUFFER_CONSTANT;
    if (copy && !host_ptr && count) return fail(QCD_E_ARG, "buffer type %d copies the host data: null host pointer", type);
    DevBuffer d;
    d.type = type;
    d.bytes = count * elem_size;
    d.host = host_ptr;
    d.live = true;
    if (type != QCD_BUFFER_LDS && d.bytes) {   // LDS: a size record only (a __local argument)
        if (hipMalloc(&d.dev, d.bytes) != hipSuccess) return fail(QCD_E_NOMEM, "buffer of %zu bytes", d.bytes);
        const hipError_t e = copy ? hipMemcpy(d.dev, host_ptr, d.bytes, hipMemcpyHostToDevice) : hipMemset(d.dev, 0, d.bytes);
        if (e != hipSuccess) {
            (void)hipFree(d.dev);
            return fail(QCD_E_HIP, "buffer set-up: %s", hipGetErrorString(e));
        }
    }
    buffers().push_back(std::move(d));
    *id = (int)buffers().size() - 1;
    return QCD_OK;
}

int qcd_buffer_write(int id)
{
    DevBuffer* d = buffer_of(id);
    if (!d) return fail(QCD_E_ARG, "no buffer %d", id);
    if (!d->dev) return d->bytes ? fail(QCD_E_STATE, "buffer %d has no device memory (LDS)", id) : QCD_OK;
    if (!d->host) return fail(QCD_E_STATE, "buffer %d was created without a host copy", id);
    HIPCHK(hipMemcpy(d->dev, d->host, d->bytes, hipMemcpyHostToDevice));
    return QCD_OK;
}

int qcd_buffer_map(int id, void** host)
{
    DevBuffer* d = buffer_of(id);
    if (!d || !host) return fail(QCD_E_ARG, "no buffer %d", id);
    if (!d->dev && d->bytes) return fail(QCD_E_STATE, "buffer %d has no device memory (LDS)", id);
    d->mirror.resize(d->bytes);
    if (d->bytes) HIPCHK(hipMemcpy(d->mirror.data(), d->dev, d->bytes, hipMemcpyDeviceToHost));
    *host = d->mirror.data();
    return QCD_OK;
}

int qcd_buffer_unmap(int id)
{
    DevBuffer* d = buffer_of(id);
    if (!d) return fail(QCD_E_ARG, "no buffer %d", id);
    if (d->mirror.size() != d->bytes) return fail(QCD_E_STATE, "buffer %d is not mapped", id);
    if (d->bytes) HIPCHK(hipMemcpy(d->dev, d->mirror.data(), d->bytes, hipMemcpyHostToDevice));
    return QCD_OK;
}

int qcd_buffer_info(int id, size_t* bytes, void** device_ptr)
{
    DevBuffer* d = buffer_of(id);
    if (!d) return fail(QCD_E_ARG, "no buffer %d", id);
    if (bytes) *bytes = d->bytes;
    if (device_ptr) *device_ptr = d->dev;
    return QCD_OK;
}

int qcd_buffer_kill(int id)
{
    DevBuffer* d = buffer_of(id);
    if (!d) return fail(QCD_E_ARG, "no buffer %d", id);
    if (d->dev) HIPCHK(hipFree(d->dev));
    *d = DevBuffer();
    return QCD_OK;
}

unsigned qcd_buffer_size_align(unsigned size, unsigned step)
{
    if (step == 0) step = 1024;   // the device's largest workgroup
    return size ? (1u + (size - 1u) / step) * step : 0u;
}

int qcd_graph_begin(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (L->graph_open) return fail(QCD_E_STATE, "a graph is already being recorded");
    if (L->profiling) return fail(QCD_E_STATE, "kernel timing waits for every kernel: no graph while it is on");
    drop_cycle_graphs(L);
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
    L->graph_open = true;
    L->graph_counter0 = L->prng_counter;
    L->graph_rng0 = L->rng_cur;
    L->cyc_hb = L->cyc_or = L->cyc_mask = 0;
    return QCD_OK;
}

int qcd_graph_end(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (!L->graph_open) return fail(QCD_E_STATE, "no graph is being recorded");
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(L->stream, &graph);
    L->graph_open = false;
    // productions count when the graph is launched, not when recorded
    L->cyc_prn = L->prng_counter - L->graph_counter0;
    L->prng_counter = L->graph_counter0;
    if (e != hipSuccess) return fail(QCD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    // fused XOR128 / PM heatbath half-steps alternate the state buffers: a
    // replayed graph must end on the buffer it starts from
    if (L->rng_cur != L->graph_rng0) {
        (void)hipGraphDestroy(graph);
        L->rng_cur = L->graph_rng0;
        return fail(QCD_E_STATE, "a graph needs an even number of heatbath half-steps with fused streams");
    }
    L->graph = graph;
    HIPCHK(hipGraphInstantiate(&L->exec, graph, nullptr, nullptr, 0));
    return QCD_OK;
}

// x-slab: the one-GPU share of a lattice cut into slabs along x (BIGLAT,
// suncl.cpp:2422 N1 = Nx + 2, UpdateEdges :4005-4062).  Half-steps update the
// interior slices only; after each one the slab's first and last interior
// slices of the updated (parity, mu) group go to the neighbours' halos.
int qcd_lattice_set_slab(qcd_lattice* L, int on)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (on) {
        const uint32_t hn = L->g.vh / (uint32_t)L->g.n1;   // links of one parity and direction in an x-slice
        if (L->g.n1 < 4) return fail(QCD_E_ARG, "an x-slab needs N1 >= 4 (two halo slices and an interior)");
        if (hn % kLinks) return fail(QCD_E_ARG, "an x-slab needs N2 N3 N4 / 2 divisible by 64 (got %u)", hn);
        if (L->tw.on) return fail(QCD_E_UNSUPPORTED, "no twisted boundary on an x-slab");
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    if (L->slab != (on != 0)) drop_cycle_graphs(L);   // a captured cycle holds the other launch shape
    L->slab = on != 0;
    return QCD_OK;
}

int qcd_lattice_profile(qcd_lattice* L, int on)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (on && !L->prof_ev[0]) {
        HIPCHK(hipEventCreate(&L->prof_ev[0]));
        HIPCHK(hipEventCreate(&L->prof_ev[1]));
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    L->profiling = on != 0;
    if (on) L->prof.clear();
    return QCD_OK;
}

int qcd_profile_count(const qcd_lattice* L, int* n)
{
    if (!L || !n) return fail(QCD_E_ARG, "null argument");
    *n = (int)L->prof.size();
    return QCD_OK;
}

int qcd_profile_entry(const qcd_lattice* L, int i, char* name, int name_len, unsigned* starts, double* mean_ms,
                      double* dev_ms)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (i < 0 || i >= (int)L->prof.size()) return fail(QCD_E_ARG, "profile entry %d out of range", i);
    const auto& e = L->prof[i];
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", e.name.c_str());
    const double m = e.sum / e.n, var = e.sumsq / e.n - m * m;
    if (starts) *starts = e.n;
    if (mean_ms) *mean_ms = m;
    if (dev_ms) *dev_ms = var > 0.0 ? sqrt(var) : 0.0;
    return QCD_OK;
}

int qcd_slab_face_bytes(const qcd_lattice* L, size_t* bytes)
{
    if (!L || !bytes) return fail(QCD_E_ARG, "null argument");
    *bytes = (size_t)L->np * (L->g.vh / (uint32_t)L->g.n1) * L->elem;
    return QCD_OK;
}

int qcd_slab_face(qcd_lattice* L, int side, int parity, int mu, void* buf, int unpack)
{
    if (!L || !buf) return fail(QCD_E_ARG, "null argument");
    if (!L->slab) return fail(QCD_E_STATE, "not an x-slab (qcd_lattice_set_slab)");
    if ((side != 0 && side != 1) || (parity != 0 && parity != 1) || mu < 0 || mu > 3)
        return fail(QCD_E_ARG, "side, parity 0 or 1 and mu 0..3");
    const size_t hn = L->g.vh / (uint32_t)L->g.n1;
    // pack: first (side 0) or last (side 1) interior slice; unpack: the halo
    // on that side
    const size_t x = unpack ? (side ? (size_t)L->g.n1 - 1 : 0) : (side ? (size_t)L->g.n1 - 2 : 1);
    char* tab = (char*)L->table;
    for (int k = 0; k < L->np; ++k) {
        char* plane = tab + ((size_t)((parity * 4 + mu) * L->np + k) * L->g.vh_pad + x * hn) * L->elem;
        char* b = (char*)buf + (size_t)k * hn * L->elem;
        HIPCHK(hipMemcpyAsync(unpack ? plane : b, unpack ? b : plane, hn * L->elem, hipMemcpyDeviceToDevice,
                              L->stream));
    }
    return QCD_OK;
}

int qcd_lattice_tune(qcd_lattice* L, int knob, int value)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    switch (knob) {
        case QCD_TUNE_XCD_MAP:
            if (value < 0 || value > 3) return fail(QCD_E_ARG, "xcd map must be 0, 1, 2 or 3");
            L->tune_xcd_map = value;
            break;
        case QCD_TUNE_BLOCKS_PER_CU:
            if (value < 0 || value > 32) return fail(QCD_E_ARG, "blocks per CU must be 0..32");
            L->tune_blocks_per_cu = value;
            break;
        case QCD_TUNE_UPDATE_SHAPE:
            if (value != 1 && value != 3) return fail(QCD_E_ARG, "update shape must be 1 or 3");
            L->tune_update_shape = value;
            break;
        case QCD_TUNE_CYCLES_PER_GRAPH:
            if (value < 1 || value > 64) return fail(QCD_E_ARG, "cycles per graph must be 1..64");
            L->tune_cycles = value;   // from the next qcd_cycle_capture
            break;
        case QCD_TUNE_SHADOW:
            if (value != 0 && value != 1) return fail(QCD_E_ARG, "shadow mode must be 0 or 1");
            if (L->batch_owned) return fail(QCD_E_STATE, "shadow mode applies to single lattices, not batched chains");
            if (value && !L->shadow) {
                if (hipMalloc(&L->shadow, L->table_bytes) != hipSuccess) {
                    L->shadow = nullptr;
                    return fail(QCD_E_NOMEM, "shadow table (%zu bytes)", L->table_bytes);
                }
            } else if (!value && L->shadow) {
                HIPCHK(hipStreamSynchronize(L->stream));
                HIPCHK(hipFree(L->shadow));
                L->shadow = nullptr;
            }
            break;
        default:
            return fail(QCD_E_ARG, "unknown tuning knob %d", knob);
    }
    return QCD_OK;
}

int qcd_lattice_get_tune(const qcd_lattice* L, int knob, int* value, int* detail)
{
    if (!L || !value) return fail(QCD_E_ARG, "null argument");
    switch (knob) {
        case QCD_TUNE_XCD_MAP: {
            UpdArgs a{};
            dispatch_order(L, L->tune_xcd_map, L->tune_update_shape == 1 && !L->tw.on, a);
            *value = a.xcd_map;
            if (detail) {   // XCD block: x-values, z-rows (modes 2 and 3)
                detail[0] = a.xcd_map >= 2 ? (int)a.ord_xs : 0;
                detail[1] = a.xcd_map >= 2 ? (int)(a.ord_zr * ((uint32_t)L->g.n3 / a.ord_ng)) : 0;
            }
            break;
        }
        case QCD_TUNE_BLOCKS_PER_CU: *value = L->tune_blocks_per_cu; break;
        case QCD_TUNE_UPDATE_SHAPE: *value = L->tw.on ? 3 : L->tune_update_shape; break;
        case QCD_TUNE_SHADOW: *value = L->shadow ? 1 : 0; break;
        case QCD_TUNE_CYCLES_PER_GRAPH:
            *value = L->tune_cycles;
            if (detail) detail[0] = L->cyc_k;   // cycles in the captured multi-cycle graph (0: none)
            break;
        default: return fail(QCD_E_ARG, "unknown tuning knob %d", knob);
    }
    return QCD_OK;
}

int qcd_lattice_table_bytes(const qcd_lattice* L, size_t* bytes)
{
    if (!L || !bytes) return fail(QCD_E_ARG, "null argument");
    *bytes = L->table_bytes;
    return QCD_OK;
}

// The reference's hot start (suncl.cpp:5234-5248): the run's streams are
// seeded with the hot runs' produce() size, then one produce() before each
// of the four direction kernels; only the first `need` quads of every
// stream feed the links, the rest of each produce() is skipped.
static int random_start(qcd_lattice* L)
{
    if (L->p.prng != QCD_PRNG_XOR128 && L->p.prng != QCD_PRNG_PM && L->p.prng != QCD_PRNG_RANLUX)
        return fail(QCD_E_UNSUPPORTED, "generator %d not available for the random start", L->p.prng);
    if (L->p.prng != QCD_PRNG_RANLUX && L->p.stream_mode != QCD_STREAM_REFERENCE)
        return fail(QCD_E_UNSUPPORTED, "the random start reads produce() buffers: it needs QCD_STREAM_REFERENCE");
    bool changed = false;
    int rc = set_sizing(L, true, &changed);
    if (rc) return rc;
    if ((rc = qcd_prng_init(L))) return rc;   // seeding and the initial produce()
    L->gid_updates = false;
    const uint32_t n = L->n_inst, S = L->s_quads;
    const uint64_t V = 2ull * L->g.vh;
    const uint64_t rowsize = (V + 63u) / 64u * 64u;
    const uint64_t used = (L->p.sun == 3) ? V + 2u * rowsize : V;   // quads one direction reads
    const uint32_t need = (uint32_t)((used + n - 1) / n);
    if (need > S) return fail(QCD_E_STATE, "internal: random start reads %u of %u quads", need, S);
    float4* buf = L->prn_buf;   // the quad buffer: prn_quads >= S quads per instance
    if (!L->quad_buffer) HIPCHK(hipMalloc((void**)&buf, sizeof(float4) * (size_t)need * n));
    const uint32_t nb = align_up(n, kBlock) / kBlock;
    for (int d = 0; d < 4 && rc == QCD_OK; ++d) {
        if (L->p.prng == QCD_PRNG_RANLUX) {
            launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, n, need, buf, L->stream);
            if (S > need) launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, n, S - need, nullptr, L->stream);
        } else if (L->p.prng == QCD_PRNG_XOR128) {
            hipLaunchKernelGGL(k_xor128_produce, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, need,
                               buf);
            hipLaunchKernelGGL(k_xor128_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, S - need);
        } else {
            hipLaunchKernelGGL(k_pm_produce, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, need, buf);
            hipLaunchKernelGGL(k_pm_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], n, S - need);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(QCD_E_HIP, "random start produce: %s", hipGetErrorString(e));
        if (rc == QCD_OK) rc = DISPATCH(L, launch_random, L, (const float4*)buf, (uint32_t)rowsize, d);
        if (rc == QCD_OK) ++L->prng_counter;
    }
    const hipError_t e = hipStreamSynchronize(L->stream);
    if (!L->quad_buffer) (void)hipFree(buf);
    if (rc == QCD_OK && e != hipSuccess) rc = fail(QCD_E_HIP, "random start: %s", hipGetErrorString(e));
    return rc;
}

int qcd_lattice_init(qcd_lattice* L, int start)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (start == QCD_START_RANDOM) return random_start(L);
    if (start != QCD_START_COLD && start != QCD_START_GID && start != QCD_START_HOT)
        return fail(QCD_E_ARG, "bad start type %d", start);
    {   // these starts run with the cold runs' produce() size; streams seeded
        // under the hot runs' size are seeded again
        bool changed = false;
        int rc = set_sizing(L, false, &changed);
        if (rc) return rc;
        if (changed && L->rng_ready && (rc = qcd_prng_init(L))) return rc;
    }
    // The GID known-answer start (INTS = 2) also switches the heatbath to the
    // GID numbers, as -D GID_UPD does in the reference (suncl.cpp:3278-3279).
    if (start == QCD_START_COLD) {
        L->gid_updates = false;
        ProfScope ps(L, "cold start (lattice_init_cold)");
        return DISPATCH(L, launch_simple, L, 0, LaunchOpt{});
    }
    if (start == QCD_START_GID) {
        L->gid_updates = true;
        ProfScope ps(L, "GID start (lattice_init_gid)");
        return DISPATCH(L, launch_simple, L, 1, LaunchOpt{});
    }
    if (start == QCD_START_HOT) {
        if (!L->rng_ready) return fail(QCD_E_STATE, "the hot start draws from the streams: call qcd_prng_init first");
        if (L->p.prng_instances)
            return fail(QCD_E_UNSUPPORTED, "the Haar start draws one stream per site pair: not with prng_instances set "
                                           "(cold, GID or the reference's random start)");
        L->gid_updates = false;
        ProfScope ps(L, "Haar start");
        const int rc = DISPATCH(L, launch_simple, L, 5, LaunchOpt{});
        // the reference produces before each of its four init kernels
        // (suncl.cpp:5236-5248)
        if (rc == QCD_OK) L->prng_counter += 4;
        return rc;
    }
    return QCD_OK;
}

int qcd_prng_init(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    // RANDSERIES 0 takes the time as seed and keeps it as the series, so
    // checkpoints can replay it (random.cpp:228-232)
    unsigned seed = L->p.randseries ? L->p.randseries : (unsigned)time(nullptr);
    L->p.randseries = seed;
    if (L->p.prng == QCD_PRNG_RANLUX) {
        // RANLUX streams are always in the reference layout: a produce()
        // fills the quad buffer before every heatbath half-step.
        std::vector<float> tab;
        srand(seed);
        ranlux_seed_table(L->n_inst, tab);
        HIPCHK(hipMemcpyAsync(L->rl_state, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, L->stream));
        launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, L->n_inst, L->s_quads, nullptr, L->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(L->stream));
        L->rng_ready = true;
        L->prng_counter = 1;
        return QCD_OK;
    }
    if (L->p.prng != QCD_PRNG_XOR128 && L->p.prng != QCD_PRNG_PM)
        return fail(QCD_E_UNSUPPORTED, "generator %d is not available", L->p.prng);
    std::vector<uint32_t> tab((size_t)4 * L->n_inst);
    srand(seed);
    // XOR128: instance 0 = XOR128_initialize_CPU (4 x rand(), random.cpp:675-681),
    // others 4 x rand() (random.cpp:653-659).  PM: every entry rand() % (2^31-1)
    // (random.cpp:867-876,892-894).
    for (size_t k = 0; k < tab.size(); ++k)
        tab[k] = (L->p.prng == QCD_PRNG_XOR128) ? (uint32_t)rand() : (uint32_t)(rand() % 2147483647);
    L->rng_cur = 0;
    HIPCHK(hipMemcpyAsync(L->rng_buf[0], tab.data(), sizeof(uint32_t) * tab.size(), hipMemcpyHostToDevice, L->stream));
    // The initial produce() of PRNG::initialize (random.cpp:276) consumes S
    // quads of every instance before the first half-step.
    uint32_t nb = align_up(L->n_inst, kBlock) / kBlock;
    if (L->p.stream_mode == QCD_STREAM_REFERENCE) {
        if (L->p.prng == QCD_PRNG_XOR128)
            hipLaunchKernelGGL(k_xor128_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[0], L->n_inst, L->s_quads);
        else
            hipLaunchKernelGGL(k_pm_skip, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[0], L->n_inst, L->s_quads);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    L->rng_ready = true;
    L->prng_counter = 1;
    return QCD_OK;
}

int qcd_lattice_import(qcd_lattice* L, const double* links)
{
    if (!L || !links) return fail(QCD_E_ARG, "null argument");
    const int N = L->p.sun;
    std::vector<unsigned char> host(L->table_bytes, 0);
    const uint32_t vh = L->g.vh, vp = L->g.vh_pad;
    const uint64_t V = 2ull * vh;
    for (uint64_t s = 0; s < V; ++s) {
        uint32_t y = (uint32_t)(s % L->g.n2);
        uint64_t r = s / L->g.n2;
        uint32_t z = (uint32_t)(r % L->g.n3);
        r /= L->g.n3;
        uint32_t t = (uint32_t)(r % L->g.n4);
        uint32_t x = (uint32_t)(r / L->g.n4);
        int par = (int)((x + y + z + t) & 1);
        uint64_t i = s >> 1;
        for (int d = 0; d < 4; ++d)
            for (int k = 0; k < L->np; ++k) {
                const double* src = links + 2 * (((size_t)s * 4 + d) * N * N + k);
                size_t off = ((size_t)((par * 4 + d) * L->np + k) * vp + i);
                if (L->elem == 8) {
                    float* dst = (float*)host.data() + 2 * off;
                    dst[0] = (float)src[0];
                    dst[1] = (float)src[1];
                } else {
                    double* dst = (double*)host.data() + 2 * off;
                    dst[0] = src[0];
                    dst[1] = src[1];
                }
            }
    }
    HIPCHK(hipMemcpyAsync(L->table, host.data(), L->table_bytes, hipMemcpyHostToDevice, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return QCD_OK;
}

int qcd_lattice_export(qcd_lattice* L, double* links)
{
    if (!L || !links) return fail(QCD_E_ARG, "null argument");
    const int N = L->p.sun;
    std::vector<unsigned char> host(L->table_bytes);
    HIPCHK(hipMemcpyAsync(host.data(), L->table, L->table_bytes, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const uint32_t vh = L->g.vh, vp = L->g.vh_pad;
    const uint64_t V = 2ull * vh;
    for (uint64_t s = 0; s < V; ++s) {
        uint32_t y = (uint32_t)(s % L->g.n2);
        uint64_t r = s / L->g.n2;
        uint32_t z = (uint32_t)(r % L->g.n3);
        r /= L->g.n3;
        uint32_t t = (uint32_t)(r % L->g.n4);
        uint32_t x = (uint32_t)(r / L->g.n4);
        int par = (int)((x + y + z + t) & 1);
        uint64_t i = s >> 1;
        for (int d = 0; d < 4; ++d) {
            double* m = links + 2 * (((size_t)s * 4 + d) * N * N);
            for (int k = 0; k < L->np; ++k) {
                size_t off = ((size_t)((par * 4 + d) * L->np + k) * vp + i);
                if (L->elem == 8) {
                    const float* src = (const float*)host.data() + 2 * off;
                    m[2 * k] = src[0];
                    m[2 * k + 1] = src[1];
                } else {
                    const double* src = (const double*)host.data() + 2 * off;
                    m[2 * k] = src[0];
                    m[2 * k + 1] = src[1];
                }
            }
            if (N == 3 && L->np == 6) {
                // row 3 = conj(row1 x row2)
                double* u = m;
                double* v = m + 6;
                double* w = m + 12;
                auto cm = [](double ar, double ai, double br, double bi, double& cr, double& ci) {
                    cr = ar * br - ai * bi;
                    ci = ar * bi + ai * br;
                };
                double p1r, p1i, p2r, p2i;
                cm(u[2], u[3], v[4], v[5], p1r, p1i);
                cm(u[4], u[5], v[2], v[3], p2r, p2i);
                w[0] = p1r - p2r;
                w[1] = -(p1i - p2i);
                cm(u[4], u[5], v[0], v[1], p1r, p1i);
                cm(u[0], u[1], v[4], v[5], p2r, p2i);
                w[2] = p1r - p2r;
                w[3] = -(p1i - p2i);
                cm(u[0], u[1], v[2], v[3], p1r, p1i);
                cm(u[2], u[3], v[0], v[1], p2r, p2i);
                w[4] = p1r - p2r;
                w[5] = -(p1i - p2i);
            }
            if (N == 2) {
                m[4] = -m[2];
                m[5] = m[3];
                m[6] = m[0];
                m[7] = -m[1];
            }
        }
    }
    return QCD_OK;
}

static int halfstep_impl(qcd_lattice* L, int parity, int mu, int kind, const LaunchOpt& o)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if ((parity != 0 && parity != 1) || mu < 0 || mu > 3) return fail(QCD_E_ARG, "bad parity/mu");
    if (kind != QCD_HEATBATH && kind != QCD_OVERRELAX) return fail(QCD_E_ARG, "bad update kind");
    return DISPATCH(L, launch_update_rn, L, parity, mu, kind, L->gid_updates ? 1 : 0, o);
}

int qcd_halfstep(qcd_lattice* L, int parity, int mu, int kind)
{
    return halfstep_impl(L, parity, mu, kind, LaunchOpt{});
}

// One sweep; with fuse_plaquette the even Y, Z and T half-steps also produce
// the plaquette record of the configuration the sweep leaves behind.
static int sweep_impl(qcd_lattice* L, int kind, bool fuse_plaquette, std::vector<hipEvent_t>* ev, int* ev_used,
                      hipEvent_t join_before_odd_t = nullptr)
{
    for (int parity = 1; parity >= 0; --parity)
        for (int mu = 0; mu < 4; ++mu) {
            // the odd T half-step is the first to write links a measurement
            // of the previous cycle (plaquette sums, Polyakov loop) reads
            if (join_before_odd_t && parity == 1 && mu == 3) HIPCHK(hipStreamWaitEvent(L->stream, join_before_odd_t, 0));
            LaunchOpt o;
            if (fuse_plaquette && parity == 0 && mu >= 1) o.meas_mode = mu;   // 1, 2, 3
            if (ev) {
                o.e0 = (*ev)[2 * (*ev_used)];
                o.e1 = (*ev)[2 * (*ev_used) + 1];
                ++*ev_used;
            }
            int rc = halfstep_impl(L, parity, mu, kind, o);
            if (rc) return rc;
        }
    return QCD_OK;
}

int qcd_sweep(qcd_lattice* L, int kind)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    return sweep_impl(L, kind, false, nullptr, nullptr);
}

int qcd_reunitarize(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    return DISPATCH(L, launch_simple, L, 2, LaunchOpt{});
}

static int measure_impl(qcd_lattice* L, int mask, bool fused_plq, const LaunchOpt& o_plq, const LaunchOpt& o_poly,
                        const LaunchOpt& o_fin)
{
    int rc = QCD_OK;
    int flags = 0;
    const int fmask = mask & (QCD_MEAS_FMUNU | QCD_MEAS_F0MU);
    if (fmask == (QCD_MEAS_FMUNU | QCD_MEAS_F0MU)) return fail(QCD_E_ARG, "F_munu: H field or E field, not both");
    if (fmask || (mask & QCD_MEAS_ACTION_SLICES)) {   // recorded beside the plaquette, from the standalone kernel
        mask |= QCD_MEAS_PLAQUETTE;
        fused_plq = false;
    }
    if ((mask & QCD_MEAS_ACTION_SLICES) && !L->site_pl) {
        // per-site sums: allocated on first use (16 bytes per site)
        if (hipMalloc((void**)&L->site_pl, sizeof(double) * 4 * (size_t)L->g.vh) != hipSuccess) {
            L->site_pl = nullptr;
            return fail(QCD_E_NOMEM, "action slices: per-site buffer");
        }
    }
    if (fused_plq) flags |= 1;
    if ((mask & QCD_MEAS_PLAQUETTE) && !fused_plq) {
        L->want_aslices = (mask & QCD_MEAS_ACTION_SLICES) != 0;
        ProfScope ps(L, "plaquette (lattice_measurement_plq)");
        rc = DISPATCH(L, launch_simple, L, 3, o_plq);
        L->want_aslices = false;
        flags |= 2;
        if (mask & QCD_MEAS_ACTION_SLICES) flags |= 64;
    }
    if (rc) return rc;
    if (fmask) {
        L->f_efield = (fmask == QCD_MEAS_F0MU) ? 1 : 0;
        ProfScope ps(L, "F_munu (lattice_measurement)");
        rc = DISPATCH(L, launch_simple, L, 7, LaunchOpt{});
        flags |= 32;
    }
    if (rc) return rc;
    if (mask & (QCD_MEAS_POLYAKOV | QCD_MEAS_POLYAKOV_SLICES)) {
        L->want_slices = (mask & QCD_MEAS_POLYAKOV_SLICES) != 0;
        ProfScope ps(L, "Polyakov loop (lattice_polyakov)");
        rc = DISPATCH(L, launch_simple, L, 4, o_poly);
        L->want_slices = false;
        flags |= 4;
        if (mask & QCD_MEAS_POLYAKOV_SLICES) flags |= 16;
    }
    if (rc) return rc;
    if (mask & QCD_MEAS_WILSON) {
        ProfScope ps(L, "Wilson loop (lattice_measurement_wilson)");
        rc = DISPATCH(L, launch_simple, L, 6, LaunchOpt{});
        flags |= 8;
    }
    if (rc || !flags) return rc;
    ProfScope ps(L, "final sums (reduce)");
    return launch_finalize(L, flags, o_fin);
}

int qcd_measure(qcd_lattice* L, int mask)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    return measure_impl(L, mask, false, LaunchOpt{}, LaunchOpt{}, LaunchOpt{});
}

// qcd_measure with every partial sum and per-site buffer in the device's
// `scratch` array instead of the lattice's own (or lazily allocated)
// buffers: no allocation, so it can run under stream capture or in loops
// that must not allocate.  Each kernel writes every slot it later reads, so
// nothing in the array needs clearing.  Calls that overlap on one GPU share
// the array and race.
int qcd_measure_noalloc(qcd_lattice* L, int mask)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (g_device < 0 || g_device >= kMaxDevices || !g_scratch[g_device])
        return fail(QCD_E_STATE, "no device scratch (qcd_device_open failed or was not called)");
    const uint32_t vs = L->g.vh * 2u / (uint32_t)L->g.n4;
    const size_t n_plq = (size_t)4 * L->upd_blocks, n_poly = (size_t)4 * L->poly_blocks;
    const size_t n_wil = L->wil_blocks, n_f = (size_t)24 * L->upd_blocks;
    const size_t n_site_pl = (mask & QCD_MEAS_ACTION_SLICES) ? (size_t)4 * L->g.vh : 0;
    const size_t n_site_tr = (mask & QCD_MEAS_POLYAKOV_SLICES) ? (size_t)2 * vs : 0;
    // every region starts on a 256-byte boundary
    auto up = [](size_t n) { return (n + 31) & ~(size_t)31; };
    const size_t need = up(n_plq) + up(n_poly) + up(n_wil) + up(n_f) + up(n_site_pl) + up(n_site_tr);
    if (need * sizeof(double) > kScratchBytes)
        return fail(QCD_E_ARG, "lattice too large for the allocation-free measurement: %zu bytes of scratch needed, "
                               "%zu available", need * sizeof(double), kScratchBytes);
    double* base = g_scratch[g_device];
    double* const saved[6] = {L->part_plq, L->part_poly, L->part_wil, L->part_f, L->site_pl, L->site_tr};
    L->part_plq = base;
    L->part_poly = L->part_plq + up(n_plq);
    L->part_wil = L->part_poly + up(n_poly);
    L->part_f = L->part_wil + up(n_wil);
    double* const pl = L->part_f + up(n_f);
    if (n_site_pl) L->site_pl = pl;
    if (n_site_tr) L->site_tr = pl + up(n_site_pl);
    const int rc = measure_impl(L, mask, false, LaunchOpt{}, LaunchOpt{}, LaunchOpt{});
    L->part_plq = saved[0];
    L->part_poly = saved[1];
    L->part_wil = saved[2];
    L->part_f = saved[3];
    L->site_pl = saved[4];
    L->site_tr = saved[5];
    return rc;
}

int qcd_read_measurements(qcd_lattice* L, double* plq, double* poly, int* count_plq, int* count_poly)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    int idx[4];
    HIPCHK(hipMemcpyAsync(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    int np = idx[0] < L->p.max_measurements ? idx[0] : L->p.max_measurements;
    int nq = idx[1] < L->p.max_measurements ? idx[1] : L->p.max_measurements;
    if (plq && np) HIPCHK(hipMemcpy(plq, L->plq_rec, sizeof(double) * 2 * np, hipMemcpyDeviceToHost));
    if (poly && nq) HIPCHK(hipMemcpy(poly, L->poly_rec, sizeof(double) * 4 * nq, hipMemcpyDeviceToHost));
    if (count_plq) *count_plq = idx[0];
    if (count_poly) *count_poly = idx[1];
    return QCD_OK;
}

int qcd_lattice_set_fmunu(qcd_lattice* L, int gen1, int gen2)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    const bool ok = (L->p.sun == 3) ? (gen1 >= 1 && gen1 <= 4 && gen2 >= 5 && gen2 <= 8)
                                    : ((gen1 == 1 || gen1 == 3) && gen2 == 2);
    if (!ok)
        return fail(QCD_E_ARG, "F_munu generators %d, %d: SU(3) takes 1..4 and 5..8, SU(2) 1 or 3 and 2", gen1, gen2);
    L->f_gen[0] = gen1;
    L->f_gen[1] = gen2;
    return QCD_OK;
}

int qcd_lattice_set_twist(qcd_lattice* L, double phi, double omega)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (L->slab && (phi != 0.0 || omega != 0.0)) return fail(QCD_E_UNSUPPORTED, "no twisted boundary on an x-slab");
    if (!std::isfinite(phi) || !std::isfinite(omega)) return fail(QCD_E_ARG, "PHI and OMEGA must be finite");
    L->tw_phi = phi;
    L->tw_omega = omega;
    Twist& t = L->tw;
    // SU(3): diag(e^{i(phi+omega)/2}, e^{i(omega-phi)/2}, e^{-i omega}); SU(2):
    // diag(e^{i phi/2}, e^{-i phi/2}), OMEGA unused (su2_matrix_memory.cl:56-88)
    const double ang[3] = {L->p.sun == 3 ? 0.5 * (phi + omega) : 0.5 * phi,
                           L->p.sun == 3 ? 0.5 * (omega - phi) : -0.5 * phi, L->p.sun == 3 ? -omega : 0.0};
    for (int r = 0; r < 3; ++r) {
        t.d[2 * r] = cos(ang[r]);
        t.d[2 * r + 1] = sin(ang[r]);
    }
    // the reference builds with TBC when PHI or OMEGA is non-zero (suncl.cpp:2415)
    t.on = (L->p.sun == 3 ? (phi != 0.0 || omega != 0.0) : phi != 0.0) ? 1 : 0;
    t.from = t.on ? L->g.vh - L->g.vh / (uint32_t)L->g.n1 : 0xffffffffu;   // first site with x = N1 - 1
    return QCD_OK;
}

int qcd_lattice_get_twist(qcd_lattice* L, double* phi, double* omega)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (phi) *phi = L->tw_phi;
    if (omega) *omega = L->tw_omega;
    return QCD_OK;
}

int qcd_read_fmunu(qcd_lattice* L, double* out, int* count)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    int idx[4];
    HIPCHK(hipMemcpyAsync(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const int n = idx[0] < L->p.max_measurements ? idx[0] : L->p.max_measurements;
    if (out && n) HIPCHK(hipMemcpy(out, L->f_rec, sizeof(double) * 12 * n, hipMemcpyDeviceToHost));
    if (count) *count = idx[0];
    return QCD_OK;
}

int qcd_read_action_slices(qcd_lattice* L, double* out, int* count)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    int idx[4];
    HIPCHK(hipMemcpyAsync(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const int n = idx[0] < L->p.max_measurements ? idx[0] : L->p.max_measurements;
    const size_t per = 2 * (size_t)(L->g.n1 + L->g.n2 + L->g.n3);
    if (out && n) HIPCHK(hipMemcpy(out, L->a_rec, sizeof(double) * per * n, hipMemcpyDeviceToHost));
    if (count) *count = idx[0];
    return QCD_OK;
}

int qcd_read_polyakov_slices(qcd_lattice* L, double* out, int* count)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    int idx[4];
    HIPCHK(hipMemcpyAsync(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const int n = idx[3] < L->p.max_measurements ? idx[3] : L->p.max_measurements;
    const size_t per = 2 * (size_t)(L->g.n1 + L->g.n2 + L->g.n3);
    if (out && n) HIPCHK(hipMemcpy(out, L->slice_rec, sizeof(double) * per * n, hipMemcpyDeviceToHost));
    if (count) *count = idx[3];
    return QCD_OK;
}

int qcd_read_wilson(qcd_lattice* L, double* w, int* count)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    int idx[4];
    HIPCHK(hipMemcpyAsync(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const int nw = idx[2] < L->p.max_measurements ? idx[2] : L->p.max_measurements;
    if (w && nw) HIPCHK(hipMemcpy(w, L->wil_rec, sizeof(double) * nw, hipMemcpyDeviceToHost));
    if (count) *count = idx[2];
    return QCD_OK;
}

int qcd_lattice_set_wilson(qcd_lattice* L, int R, int T)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (R < 1 || T < 1) return fail(QCD_E_ARG, "Wilson loop extents must be >= 1 (got %d x %d)", R, T);
    L->wil_R = R;
    L->wil_T = T;
    return QCD_OK;
}

int qcd_reset_measurements(qcd_lattice* L)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipMemset(L->rec_index, 0, sizeof(int) * 4));
    HIPCHK(hipDeviceSynchronize());
    return QCD_OK;
}

// Whether a cycle's plaquette rides on its last sweep.  F_munu needs the
// standalone plaquette kernel, and so does a twisted boundary (the fused sum
// takes the updated link as stored, the twisted Y links are read scaled).
static bool cycle_fuses(const qcd_lattice* L, int mask, int total)
{
    return (mask & QCD_MEAS_PLAQUETTE) && total > 0 &&
           !(mask & (QCD_MEAS_FMUNU | QCD_MEAS_F0MU | QCD_MEAS_ACTION_SLICES)) && !L->tw.on && !L->slab;
}

// A measurement cycle: n_hb heatbath sweeps, n_or overrelaxation sweeps,
// then the measurements.  The plaquette rides on the last sweep's even
// Y/Z/T half-steps when there is a sweep; the Polyakov loop is its own kernel.
// Whether a cycle's measurement can run beside the next cycle's first three
// half-steps (odd X, Y, Z): it may read only T links and its own partial
// sums -- the plaquette fused into the last sweep, the Polyakov loop -- so
// neither it nor those half-steps touch what the other writes.
static bool cycle_overlaps(const qcd_lattice* L, int mask, int total)
{
    return total > 0 && cycle_fuses(L, mask, total) && !(mask & ~(QCD_MEAS_PLAQUETTE | QCD_MEAS_POLYAKOV));
}

static int enqueue_cycle(qcd_lattice* L, int n_hb, int n_or, int mask, std::vector<hipEvent_t>* ev, int* ev_used,
                         bool join_prev = false, bool fork_meas = false)
{
    const int total = n_hb + n_or;
    const bool fuse = cycle_fuses(L, mask, total);
    for (int k = 0; k < total; ++k) {
        int rc = sweep_impl(L, k < n_hb ? QCD_HEATBATH : QCD_OVERRELAX, fuse && k == total - 1, ev, ev_used,
                            join_prev && k == 0 ? L->ev_join : nullptr);
        if (rc) return rc;
    }
    if (!mask) return QCD_OK;
    if (fork_meas) {   // on the second stream, joined by the next cycle's odd T half-step
        HIPCHK(hipEventRecord(L->ev_fork, L->stream));
        HIPCHK(hipStreamWaitEvent(L->side, L->ev_fork, 0));
        hipStream_t main_stream = L->stream;
        L->stream = L->side;
        const int rc = measure_impl(L, mask, fuse, LaunchOpt{}, LaunchOpt{}, LaunchOpt{});
        L->stream = main_stream;
        if (rc) return rc;
        HIPCHK(hipEventRecord(L->ev_join, L->side));
        return QCD_OK;
    }
    LaunchOpt op, oq, of;
    if (ev) {
        LaunchOpt* slots[3] = {&op, &oq, &of};
        const bool want[3] = {(mask & QCD_MEAS_PLAQUETTE) && !fuse, (mask & QCD_MEAS_POLYAKOV) != 0, true};
        for (int k = 0; k < 3; ++k)
            if (want[k]) {
                slots[k]->e0 = (*ev)[2 * (*ev_used)];
                slots[k]->e1 = (*ev)[2 * (*ev_used) + 1];
                ++*ev_used;
            }
    }
    return measure_impl(L, mask, fuse, op, oq, of);
}

int qcd_cycle_capture(qcd_lattice* L, int n_hb, int n_or, int mask)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (L->profiling) return fail(QCD_E_STATE, "kernel timing waits for every kernel: run cycles eagerly while it is on");
    if (n_hb < 0 || n_or < 0) return fail(QCD_E_ARG, "negative sweep count");
    drop_cycle_graphs(L);
    HIPCHK(hipStreamSynchronize(L->stream));
    // K = QCD_TUNE_CYCLES_PER_GRAPH consecutive cycles as one graph, each
    // cycle's measurement beside the next cycle's first half-steps (replays
    // of whole K-blocks); the single cycle below covers the rest
    const int K = cycle_overlaps(L, mask, n_hb + n_or) ? L->tune_cycles : 1;
    if (K > 1) {
        if (!L->side) HIPCHK(hipStreamCreateWithFlags(&L->side, hipStreamNonBlocking));
        if (!L->ev_fork) HIPCHK(hipEventCreateWithFlags(&L->ev_fork, hipEventDisableTiming));
        if (!L->ev_join) HIPCHK(hipEventCreateWithFlags(&L->ev_join, hipEventDisableTiming));
        HIPCHK(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
        const uint32_t c0 = L->prng_counter;
        int rk = QCD_OK;
        for (int j = 0; j < K && !rk; ++j) rk = enqueue_cycle(L, n_hb, n_or, mask, nullptr, nullptr, j > 0, j < K - 1);
        L->prng_counter = c0;
        hipGraph_t gk = nullptr;
        const hipError_t ek = hipStreamEndCapture(L->stream, &gk);
        if (rk) {
            if (gk) (void)hipGraphDestroy(gk);
            return rk;
        }
        if (ek != hipSuccess) return fail(QCD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ek));
        L->graph_k = gk;
        HIPCHK(hipGraphInstantiate(&L->exec_k, gk, nullptr, nullptr, 0));
        L->cyc_k = K;
    }
    HIPCHK(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
    const uint32_t counter0 = L->prng_counter;
    int rc = enqueue_cycle(L, n_hb, n_or, mask, nullptr, nullptr);
    // capture runs nothing: the productions count when the graph is launched
    L->cyc_prn = L->prng_counter - counter0;
    L->prng_counter = counter0;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(L->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(QCD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    L->graph = graph;
    HIPCHK(hipGraphInstantiate(&L->exec, graph, nullptr, nullptr, 0));
    L->cyc_hb = n_hb;
    L->cyc_or = n_or;
    L->cyc_mask = mask;
    return QCD_OK;
}

int qcd_cycle_run(qcd_lattice* L, int n_hb, int n_or, int mask)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (n_hb < 0 || n_or < 0) return fail(QCD_E_ARG, "negative sweep count");
    return enqueue_cycle(L, n_hb, n_or, mask, nullptr, nullptr);
}

// `repeats` cycles: whole blocks of the multi-cycle graph, then single ones.
static int launch_cycles(qcd_lattice* L, int repeats)
{
    int r = repeats;
    if (L->exec_k && L->cyc_k > 1)
        for (; r >= L->cyc_k; r -= L->cyc_k) {
            HIPCHK(hipGraphLaunch(L->exec_k, L->stream));
            L->prng_counter += L->cyc_prn * (uint32_t)L->cyc_k;
        }
    for (; r > 0; --r) {
        HIPCHK(hipGraphLaunch(L->exec, L->stream));
        L->prng_counter += L->cyc_prn;
    }
    return QCD_OK;
}

int qcd_cycle_launch(qcd_lattice* L, int repeats)
{
    if (!L || !L->exec) return fail(QCD_E_STATE, "no captured cycle");
    return launch_cycles(L, repeats);
}

int qcd_cycle_time(qcd_lattice* L, int repeats, double* ms)
{
    if (!L || !L->exec) return fail(QCD_E_STATE, "no captured cycle");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipEventRecord(a, L->stream));
    if (int rc = launch_cycles(L, repeats)) return rc;
    HIPCHK(hipEventRecord(b, L->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = f;
    return QCD_OK;
}

// Eager replay of the captured cycle with every kernel launched through
// hipExtLaunchKernelGGL, which records a HIP event pair at the start and end
// of the kernel itself on the lattice stream: per-kernel durations without
// launch gaps.  Reports means per update kind and per cycle for the
// measurement kernels.
int qcd_cycle_profile(qcd_lattice* L, int repeats, double* hb_us, double* or_us, double* meas_us, double* total_ms)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    const int n_hb = L->cyc_hb, n_or = L->cyc_or, mask = L->cyc_mask;
    const int per = 8 * (n_hb + n_or) + 3;
    std::vector<hipEvent_t> ev((size_t)2 * per);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    double t_hb = 0.0, t_or = 0.0, t_meas = 0.0, t_tot = 0.0;
    int n_fused_seen = 0;
    HIPCHK(hipStreamSynchronize(L->stream));
    for (int r = 0; r < repeats; ++r) {
        int used = 0;
        HIPCHK(hipEventRecord(t0, L->stream));
        int rc = enqueue_cycle(L, n_hb, n_or, mask, &ev, &used);
        if (rc) return rc;
        HIPCHK(hipEventRecord(t1, L->stream));
        HIPCHK(hipEventSynchronize(t1));
        const int total = n_hb + n_or;
        const bool fuse = cycle_fuses(L, mask, total);
        for (int j = 0; j < used; ++j) {
            float f = 0.f;
            HIPCHK(hipEventElapsedTime(&f, ev[2 * j], ev[2 * j + 1]));
            const int sw = j / 8, k = j % 8;
            // the last sweep's even Y/Z/T half-steps carry the plaquette
            const bool fused_launch = fuse && sw == total - 1 && k >= 5;
            if (j >= 8 * total || fused_launch) t_meas += f;
            else if (sw < n_hb) t_hb += f;
            else t_or += f;
            if (fused_launch) ++n_fused_seen;
        }
        float f = 0.f;
        HIPCHK(hipEventElapsedTime(&f, t0, t1));
        t_tot += f;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    // Means over the plain update launches; the three half-steps that carry
    // the fused plaquette are counted as measurement time.
    const int per_rep_fused = repeats ? n_fused_seen / repeats : 0;
    const int hb_launches = 8 * n_hb - (n_or == 0 ? per_rep_fused : 0);
    const int or_launches = 8 * n_or - (n_or > 0 ? per_rep_fused : 0);
    if (hb_us) *hb_us = hb_launches > 0 ? 1000.0 * t_hb / ((double)repeats * hb_launches) : 0.0;
    if (or_us) *or_us = or_launches > 0 ? 1000.0 * t_or / ((double)repeats * or_launches) : 0.0;
    if (meas_us) *meas_us = mask ? 1000.0 * t_meas / repeats : 0.0;
    if (total_ms) *total_ms = t_tot;
    return QCD_OK;
}

int qcd_prng_produce_host(int prng, int ranlux_luxury, unsigned n, unsigned samples, void* table, float* quads)
{
    if (!table || !quads || n == 0) return fail(QCD_E_ARG, "null argument");
    if (prng != QCD_PRNG_XOR128 && prng != QCD_PRNG_PM && prng != QCD_PRNG_RANLUX)
        return fail(QCD_E_UNSUPPORTED, "generator %d not available", prng);
    if (prng == QCD_PRNG_RANLUX && (ranlux_luxury < 0 || ranlux_luxury > 4))
        return fail(QCD_E_ARG, "luxury level must be 0..4");
    int rc = ensure_device();
    if (rc) return rc;
    if (prng == QCD_PRNG_RANLUX) {
        // table: 7 float4 per instance, table[(q * n + g) * 4 + c]
        float4* ds = nullptr;
        float4* dq = nullptr;
        const size_t sbytes = sizeof(float4) * 7 * (size_t)n, qbytes = sizeof(float4) * (size_t)n * samples;
        HIPCHK(hipMalloc((void**)&ds, sbytes));
        if (hipMalloc((void**)&dq, qbytes ? qbytes : 16) != hipSuccess) {
            (void)hipFree(ds);
            return fail(QCD_E_NOMEM, "quads buffer");
        }
        hipError_t e = hipMemcpy(ds, table, sbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_ranlux_produce(ranlux_luxury, ds, n, samples, dq, 0);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess && qbytes) e = hipMemcpy(quads, dq, qbytes, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(table, ds, sbytes, hipMemcpyDeviceToHost);
        (void)hipFree(ds);
        (void)hipFree(dq);
        if (e != hipSuccess) return fail(QCD_E_HIP, "ranlux produce: %s", hipGetErrorString(e));
        return QCD_OK;
    }
    uint4* dt = nullptr;
    float4* dq = nullptr;
    size_t qbytes = sizeof(float4) * (size_t)n * samples;
    HIPCHK(hipMalloc((void**)&dt, sizeof(uint4) * n));
    if (hipMalloc((void**)&dq, qbytes) != hipSuccess) {
        (void)hipFree(dt);
        return fail(QCD_E_NOMEM, "quads buffer");
    }
    HIPCHK(hipMemcpy(dt, table, sizeof(uint4) * n, hipMemcpyHostToDevice));
    uint32_t nb = align_up(n, kBlock) / kBlock;
    if (prng == QCD_PRNG_XOR128) hipLaunchKernelGGL(k_xor128_produce, dim3(nb), dim3(kBlock), 0, 0, dt, n, samples, dq);
    else hipLaunchKernelGGL(k_pm_produce, dim3(nb), dim3(kBlock), 0, 0, dt, n, samples, dq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(quads, dq, qbytes, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(table, dt, sizeof(uint4) * n, hipMemcpyDeviceToHost));
    (void)hipFree(dt);
    (void)hipFree(dq);
    return QCD_OK;
}

// ---------------------------------------------------------------- checkpoints
// model::lattice_save_state (suncl.cpp:1857-1901) keeps the table and the
// measurement buffers and resumes the generators by replaying PRNG_counter
// productions from the seed (suncl.cpp:5349-5351).  The state image below
// holds everything later updates and measurements read: the table as stored
// on the device, the current generator states, the records and their
// indices, the production counter and the start flags.  A lattice restored
// from it continues bit-identically.

namespace {

constexpr char kStateMagic[8] = {'Q', 'C', 'D', 'S', 'T', 'A', 'T', '2'};
constexpr size_t kStateHead = 136;   // bytes of the head qcd_lattice_state_save writes

struct ByteWriter {
    unsigned char* p;
    void u32(uint32_t v) { memcpy(p, &v, 4); p += 4; }
    void u64(uint64_t v) { memcpy(p, &v, 8); p += 8; }
    void raw(const void* s, size_t n) { memcpy(p, s, n); p += n; }
};
struct ByteReader {
    const unsigned char* p;
    uint32_t u32() { uint32_t v; memcpy(&v, p, 4); p += 4; return v; }
    uint64_t u64() { uint64_t v; memcpy(&v, p, 8); p += 8; return v; }
};

size_t rng_state_bytes(const qcd_lattice* L)
{
    return (L->p.prng == QCD_PRNG_RANLUX ? 7u : 1u) * sizeof(uint4) * (size_t)L->n_inst;
}

size_t state_image_bytes(const qcd_lattice* L)
{
    return kStateHead + L->table_bytes + rng_state_bytes(L) + sizeof(double) * 7 * (size_t)L->p.max_measurements +
           2 * sizeof(double) * slice_rec_doubles(L) + sizeof(double) * 12 * (size_t)L->p.max_measurements;
}

uint64_t f64_bits(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    return b;
}

// The parameter fields in image order, with their names for error messages.
void param_fields(const qcd_params& p, uint64_t v[16])
{
    const uint64_t f[16] = {(uint64_t)p.sun,        (uint64_t)p.n[0],          (uint64_t)p.n[1],
                            (uint64_t)p.n[2],       (uint64_t)p.n[3],          (uint64_t)p.precision,
                            (uint64_t)p.storage_reals, (uint64_t)p.nhit,       (uint64_t)p.nhitpar,
                            (uint64_t)p.prng,       (uint64_t)p.ranlux_luxury, (uint64_t)p.stream_mode,
                            (uint64_t)p.randseries, (uint64_t)p.max_measurements, f64_bits(p.beta), 0};
    memcpy(v, f, sizeof(f));
}
const char* const kParamNames[15] = {"sun",  "n1",      "n2",         "n3",          "n4",
                                     "precision", "storage_reals", "nhit", "nhitpar", "prng",
                                     "ranlux_luxury", "stream_mode", "randseries", "max_measurements", "beta"};

// Columns of k applications of the GF(2) map whose column b is cols[b].
void gf2_apply(const uint4* m, const uint4& v, uint4& out)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint4 o = make_uint4(0u, 0u, 0u, 0u);
    for (int b = 0; b < 128; ++b)
        if ((w[b >> 5] >> (b & 31)) & 1u) {
            o.x ^= m[b].x;
            o.y ^= m[b].y;
            o.z ^= m[b].z;
            o.w ^= m[b].w;
        }
    out = o;
}
void gf2_power(const uint4* cols, uint64_t k, uint4* out)
{
    std::vector<uint4> r(128), b(cols, cols + 128), t(128);
    for (int j = 0; j < 128; ++j) {
        r[j] = make_uint4(0u, 0u, 0u, 0u);
        ((uint32_t*)&r[j])[j >> 5] = 1u << (j & 31);
    }
    while (k) {
        if (k & 1) {
            for (int j = 0; j < 128; ++j) gf2_apply(b.data(), r[j], t[j]);
            r.swap(t);
        }
        k >>= 1;
        if (k) {
            for (int j = 0; j < 128; ++j) gf2_apply(b.data(), b[j], t[j]);
            b.swap(t);
        }
    }
    std::copy(r.begin(), r.end(), out);
}

}  // namespace

int qcd_lattice_get_params(const qcd_lattice* L, qcd_params* out)
{
    if (!L || !out) return fail(QCD_E_ARG, "null argument");
    *out = L->p;
    return QCD_OK;
}

int qcd_prng_counter(const qcd_lattice* L, unsigned* counter)
{
    if (!L || !counter) return fail(QCD_E_ARG, "null argument");
    *counter = L->prng_counter;
    return QCD_OK;
}

int qcd_lattice_state_bytes(const qcd_lattice* L, size_t* bytes)
{
    if (!L || !bytes) return fail(QCD_E_ARG, "null argument");
    *bytes = state_image_bytes(L);
    return QCD_OK;
}

int qcd_lattice_state_save(qcd_lattice* L, void* buf, size_t bytes)
{
    if (!L || !buf) return fail(QCD_E_ARG, "null argument");
    if (bytes != state_image_bytes(L))
        return fail(QCD_E_ARG, "state buffer of %zu bytes, the lattice needs %zu", bytes, state_image_bytes(L));
    HIPCHK(hipStreamSynchronize(L->stream));
    int idx[4];
    HIPCHK(hipMemcpy(idx, L->rec_index, sizeof(idx), hipMemcpyDeviceToHost));
    ByteWriter w{(unsigned char*)buf};
    w.raw(kStateMagic, 8);
    w.u32((uint32_t)kStateHead);
    uint64_t f[16];
    param_fields(L->p, f);
    for (int k = 0; k < 14; ++k) w.u32((uint32_t)f[k]);
    w.u64(f[14]);
    w.u32(L->prng_counter);
    w.u32((L->rng_ready ? 1u : 0u) | (L->hot_run ? 2u : 0u));   // start flags
    w.u32(L->gid_updates ? 1u : 0u);
    w.u32((uint32_t)L->wil_R);
    w.u32((uint32_t)L->wil_T);
    for (int k = 0; k < 4; ++k) w.u32((uint32_t)idx[k]);
    w.u64(L->table_bytes);
    w.u64(rng_state_bytes(L));
    w.u32(L->p.prng_instances);
    w.u32(0u);
    if ((size_t)(w.p - (unsigned char*)buf) != kStateHead) return fail(QCD_E_STATE, "internal: state head size");
    HIPCHK(hipMemcpy(w.p, L->table, L->table_bytes, hipMemcpyDeviceToHost));
    w.p += L->table_bytes;
    const void* rng = (L->p.prng == QCD_PRNG_RANLUX) ? (const void*)L->rl_state : (const void*)L->rng_buf[L->rng_cur];
    HIPCHK(hipMemcpy(w.p, rng, rng_state_bytes(L), hipMemcpyDeviceToHost));
    w.p += rng_state_bytes(L);
    const size_t m = (size_t)L->p.max_measurements;
    HIPCHK(hipMemcpy(w.p, L->plq_rec, sizeof(double) * 2 * m, hipMemcpyDeviceToHost));
    w.p += sizeof(double) * 2 * m;
    HIPCHK(hipMemcpy(w.p, L->poly_rec, sizeof(double) * 4 * m, hipMemcpyDeviceToHost));
    w.p += sizeof(double) * 4 * m;
    HIPCHK(hipMemcpy(w.p, L->wil_rec, sizeof(double) * m, hipMemcpyDeviceToHost));
    w.p += sizeof(double) * m;
    HIPCHK(hipMemcpy(w.p, L->slice_rec, sizeof(double) * slice_rec_doubles(L), hipMemcpyDeviceToHost));
    w.p += sizeof(double) * slice_rec_doubles(L);
    HIPCHK(hipMemcpy(w.p, L->f_rec, sizeof(double) * 12 * m, hipMemcpyDeviceToHost));
    w.p += sizeof(double) * 12 * m;
    HIPCHK(hipMemcpy(w.p, L->a_rec, sizeof(double) * slice_rec_doubles(L), hipMemcpyDeviceToHost));
    return QCD_OK;
}

int qcd_lattice_state_load(qcd_lattice* L, const void* buf, size_t bytes)
{
    if (!L || !buf) return fail(QCD_E_ARG, "null argument");
    if (bytes != state_image_bytes(L))
        return fail(QCD_E_ARG, "state image of %zu bytes, this lattice needs %zu", bytes, state_image_bytes(L));
    ByteReader r{(const unsigned char*)buf};
    if (memcmp(r.p, kStateMagic, 8) != 0) return fail(QCD_E_ARG, "not a lattice state image");
    r.p += 8;
    if (r.u32() != (uint32_t)kStateHead) return fail(QCD_E_ARG, "state image head of another layout");
    uint64_t f[16];
    param_fields(L->p, f);
    for (int k = 0; k < 15; ++k) {
        const uint64_t v = (k < 14) ? r.u32() : r.u64();
        if (v != (k < 14 ? (uint64_t)(uint32_t)f[k] : f[k]))
            return fail(QCD_E_ARG, "state image has another %s than this lattice", kParamNames[k]);
    }
    const uint32_t counter = r.u32(), ready = r.u32(), gid = r.u32(), wr = r.u32(), wt = r.u32();
    int idx[4];
    for (int k = 0; k < 4; ++k) idx[k] = (int)r.u32();
    if (r.u64() != L->table_bytes || r.u64() != rng_state_bytes(L))
        return fail(QCD_E_ARG, "state image sizes do not match this lattice");
    if (r.u32() != L->p.prng_instances) return fail(QCD_E_ARG, "state image has another prng_instances than this lattice");
    r.p += 4;
    if (ready & ~3u) return fail(QCD_E_ARG, "state image with unknown start flags %#x", ready);
    bool changed = false;
    int rc = set_sizing(L, (ready & 2u) != 0, &changed);   // the run's produce() size
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(L->stream));
    HIPCHK(hipMemcpy(L->table, r.p, L->table_bytes, hipMemcpyHostToDevice));
    r.p += L->table_bytes;
    void* rng = (L->p.prng == QCD_PRNG_RANLUX) ? (void*)L->rl_state : (void*)L->rng_buf[L->rng_cur];
    HIPCHK(hipMemcpy(rng, r.p, rng_state_bytes(L), hipMemcpyHostToDevice));
    r.p += rng_state_bytes(L);
    const size_t m = (size_t)L->p.max_measurements;
    HIPCHK(hipMemcpy(L->plq_rec, r.p, sizeof(double) * 2 * m, hipMemcpyHostToDevice));
    r.p += sizeof(double) * 2 * m;
    HIPCHK(hipMemcpy(L->poly_rec, r.p, sizeof(double) * 4 * m, hipMemcpyHostToDevice));
    r.p += sizeof(double) * 4 * m;
    HIPCHK(hipMemcpy(L->wil_rec, r.p, sizeof(double) * m, hipMemcpyHostToDevice));
    r.p += sizeof(double) * m;
    HIPCHK(hipMemcpy(L->slice_rec, r.p, sizeof(double) * slice_rec_doubles(L), hipMemcpyHostToDevice));
    r.p += sizeof(double) * slice_rec_doubles(L);
    HIPCHK(hipMemcpy(L->f_rec, r.p, sizeof(double) * 12 * m, hipMemcpyHostToDevice));
    r.p += sizeof(double) * 12 * m;
    HIPCHK(hipMemcpy(L->a_rec, r.p, sizeof(double) * slice_rec_doubles(L), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->rec_index, idx, sizeof(idx), hipMemcpyHostToDevice));
    L->prng_counter = counter;
    L->rng_ready = (ready & 1u) != 0;
    L->gid_updates = gid != 0;
    L->wil_R = (int)wr;
    L->wil_T = (int)wt;
    return QCD_OK;
}

int qcd_write_measurements(qcd_lattice* L, const double* plq, const double* poly, const double* wil, int count_plq,
                           int count_poly, int count_wil)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (count_plq < 0 || count_poly < 0 || count_wil < 0) return fail(QCD_E_ARG, "negative record count");
    if ((count_plq && !plq) || (count_poly && !poly) || (count_wil && !wil))
        return fail(QCD_E_ARG, "null record array with a non-zero count");
    const int m = L->p.max_measurements;
    HIPCHK(hipStreamSynchronize(L->stream));
    const int np = count_plq < m ? count_plq : m, nq = count_poly < m ? count_poly : m, nw = count_wil < m ? count_wil : m;
    if (np) HIPCHK(hipMemcpy(L->plq_rec, plq, sizeof(double) * 2 * np, hipMemcpyHostToDevice));
    if (nq) HIPCHK(hipMemcpy(L->poly_rec, poly, sizeof(double) * 4 * nq, hipMemcpyHostToDevice));
    if (nw) HIPCHK(hipMemcpy(L->wil_rec, wil, sizeof(double) * nw, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(L->f_rec, 0, sizeof(double) * 12 * m));   // restored records carry no F_munu
    HIPCHK(hipMemset(L->a_rec, 0, sizeof(double) * slice_rec_doubles(L)));   // nor action slices
    const int idx[4] = {count_plq, count_poly, count_wil, 0};
    HIPCHK(hipMemcpy(L->rec_index, idx, sizeof(idx), hipMemcpyHostToDevice));
    return QCD_OK;
}

int qcd_prng_replay(qcd_lattice* L, unsigned counter, int start)
{
    if (!L) return fail(QCD_E_ARG, "null lattice");
    if (start != QCD_START_HOT && start != QCD_START_COLD && start != QCD_START_GID && start != QCD_START_RANDOM)
        return fail(QCD_E_ARG, "bad start type %d", start);
    if (start == QCD_START_RANDOM && L->p.prng != QCD_PRNG_RANLUX && L->p.stream_mode != QCD_STREAM_REFERENCE)
        return fail(QCD_E_UNSUPPORTED, "the random start reads produce() buffers: it needs QCD_STREAM_REFERENCE");
    if (start == QCD_START_HOT && L->p.prng_instances)
        return fail(QCD_E_UNSUPPORTED, "no Haar start with prng_instances set");
    bool changed = false;
    int rc = set_sizing(L, start == QCD_START_RANDOM, &changed);
    if (rc) return rc;
    if (counter == 0) {
        L->rng_ready = false;
        L->prng_counter = 0;
        return QCD_OK;
    }
    rc = qcd_prng_init(L);   // counter 1: the initial produce
    if (rc) return rc;
    uint32_t left = counter - 1;
    // a random start's four productions are ordinary ones of the hot runs' size
    if (start == QCD_START_RANDOM && left > 0 && left < 4)
        return fail(QCD_E_ARG, "counter %u cannot follow a hot start (it takes 4 productions)", counter);
    const bool ranlux = (L->p.prng == QCD_PRNG_RANLUX);
    const uint32_t nb = align_up(L->n_inst, kBlock) / kBlock;
    if (start == QCD_START_HOT && left > 0) {
        // the hot start's draw: 8 links x (3 or 1) quads per stream
        if (left < 4) return fail(QCD_E_ARG, "counter %u cannot follow a hot start (it takes 4 productions)", counter);
        const uint32_t q = (L->p.sun == 3) ? 24u : 8u;
        if (ranlux) {
            launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, L->n_inst, q, nullptr, L->stream);
        } else {
            const uint32_t nv = L->g.vh, nbv = align_up(nv, kBlock) / kBlock;
            if (L->p.prng == QCD_PRNG_XOR128)
                hipLaunchKernelGGL(k_xor128_skip, dim3(nbv), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], nv, q);
            else
                hipLaunchKernelGGL(k_pm_skip, dim3(nbv), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur], nv, q);
        }
        HIPCHK(hipGetLastError());
        left -= 4;
    }
    if (left > 0) {
        if (ranlux) {
            for (uint32_t k = 0; k < left; ++k)
                launch_ranlux_produce(L->p.ranlux_luxury, L->rl_state, L->n_inst, L->s_quads, nullptr, L->stream);
            HIPCHK(hipGetLastError());
        } else if (L->p.stream_mode != QCD_STREAM_REFERENCE) {
            return fail(QCD_E_UNSUPPORTED,
                        "fast-layout streams advance by the numbers consumed: they resume from a state image only");
        } else if (L->p.prng == QCD_PRNG_XOR128) {
            uint4 cols[128], pw[128];
            HIPCHK(hipMemcpy(cols, L->jump_cols, sizeof(cols), hipMemcpyDeviceToHost));
            gf2_power(cols, left, pw);
            uint4* d = nullptr;
            HIPCHK(hipMalloc((void**)&d, sizeof(pw)));
            if (hipMemcpy(d, pw, sizeof(pw), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(d);
                return fail(QCD_E_HIP, "jump matrix upload failed");
            }
            hipLaunchKernelGGL(k_xor128_jump_apply, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur],
                               L->n_inst, (const uint4*)d);
            const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(L->stream);
            (void)hipFree(d);
            if (e1 != hipSuccess || e2 != hipSuccess)
                return fail(QCD_E_HIP, "stream jump: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        } else {
            uint32_t ak = 1u, b = L->pm_ak;
            for (uint32_t k = left; k; k >>= 1) {
                if (k & 1u) ak = pm_mulmod(ak, b);
                b = pm_mulmod(b, b);
            }
            hipLaunchKernelGGL(k_pm_jump_apply, dim3(nb), dim3(kBlock), 0, L->stream, L->rng_buf[L->rng_cur],
                               L->n_inst, ak);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipStreamSynchronize(L->stream));
    L->prng_counter = counter;
    return QCD_OK;
}

}  // extern "C"

// ============================================================================
// Batched ensembles (host)
// ============================================================================
// A batch owns independent lattices of possibly different extents and beta
// (each a full qcd_lattice with its own table, streams and records) and
// updates all of them with one launch per half-step and bucket.  Buckets are
// the chains of equal spatial extent; each bucket has its own launch and its
// own kernel shape per update kind, chosen by qcd_batch_autotune.
struct qcd_batch {
    qcd_params p;                          // common parameters (sizes, beta, seed per chain)
    std::vector<qcd_lattice*> chains;
    hipStream_t stream = nullptr;
    ChainDesc* d_desc = nullptr;
    uint2* d_map = nullptr;                // all work maps, concatenated
    // Unit placements over the 8 XCDs (hardware workgroup b runs on XCD
    // b mod 8): 0 each chain whole on one XCD, its units in index order;
    // 1 the same with each chain's units in cross-section order (x fastest,
    // then t, then z), so a unit's x and t neighbours go out close to it;
    // 2 every chain cut into 8 contiguous slabs of units, one per XCD.
    static constexpr int kPlacements = 3;
    struct Bucket {
        std::vector<int> members;
        uint32_t off3[kPlacements] = {}, len3[kPlacements] = {};   // three-wavefront maps: one entry per workgroup
        uint32_t off1[kPlacements] = {}, len1[kPlacements] = {};   // one-wavefront maps: one entry per wavefront
        int shape[2] = {3, 3};             // [heatbath, overrelaxation]
        int place[2] = {0, 0};
    };
    std::vector<Bucket> buckets;
    std::vector<uint2> h_map;              // host copy of d_map
    // Merged maps (merge on): per update kind [heatbath, overrelaxation] and
    // kernel shape [3, 1], the chosen maps of every bucket with that shape,
    // concatenated, so one launch per (parity, mu) and shape updates all
    // chains.  Every bucket map holds a multiple of 8 workgroups, so a
    // workgroup keeps its XCD (b mod 8) in the concatenation.
    bool merge = true;
    uint2* d_mmap = nullptr;
    std::vector<uint2*> mmap_old;          // earlier merged maps: a captured cycle may still launch them
    uint32_t moff[2][2] = {}, mlen[2][2] = {};
    uint32_t off_units = 0, len_units = 0; // every (chain, unit): RANLUX produce
    uint32_t off_poly = 0, len_poly = 0;   // every (chain, Polyakov workgroup)
    int rng_cur = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

UpdArgs chain_base_args(const qcd_lattice* L)
{
    UpdArgs a{};
    a.table = L->table;
    a.src = L->table;
    a.g = L->g;
    a.beta = L->p.beta / L->p.sun;
    a.nhit = L->p.nhit;
    a.sites = 2u * L->g.vh;
    a.s_quads = (L->p.stream_mode == QCD_STREAM_REFERENCE) ? L->s_quads : 0u;
    a.partials = L->part_fused;
    a.xcd_map = 0;
    a.ord_ng = a.ord_nt = a.ord_xs = a.ord_ax = a.ord_zr = 1;
    a.jump = L->jump_cols;
    a.pm_ak = L->pm_ak;
    a.prn = L->prn_buf;
    a.prn_stride = L->n_streams;
    return a;
}

FinArgs chain_fin_args(const qcd_lattice* L)
{
    FinArgs f{};
    f.fused = L->part_fused;
    f.n_fused = L->nb_fused;
    f.plq = L->part_plq;
    f.n_plq = 2u * L->upd_blocks;
    f.poly = L->part_poly;
    f.n_poly = L->poly_blocks;
    f.wil = L->part_wil;
    f.n_wil = L->wil_blocks;
    f.plq_rec = L->plq_rec;
    f.poly_rec = L->poly_rec;
    f.wil_rec = L->wil_rec;
    f.rec_index = L->rec_index;
    f.capacity = L->p.max_measurements;
    f.f_rec = L->f_rec;
    f.a_rec = L->a_rec;
    f.a_n = L->g.n1 + L->g.n2 + L->g.n3;
    return f;
}

// The 64-link units of chain L in cross-section order: sorted by the z,
// t and x coordinates of their first link (z slowest, x fastest).
std::vector<uint32_t> cross_section_units(const qcd_lattice* L)
{
    const uint32_t nu = L->nb_fused;
    std::vector<uint64_t> key(nu);
    for (uint32_t u = 0; u < nu; ++u) {
        const Site st = site_of(L->g, u * kLinks, 0);
        key[u] = (((uint64_t)st.z * (uint64_t)L->g.n4 + (uint64_t)st.t) * (uint64_t)L->g.n1 + (uint64_t)st.x) << 32 | u;
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> out(nu);
    for (uint32_t u = 0; u < nu; ++u) out[u] = (uint32_t)(key[u] & 0xffffffffu);
    return out;
}

// Work maps of every bucket and placement.  Placements 0 and 1: chains of a
// bucket go to the XCD with the least work so far (largest first); each
// XCD's list holds its chains' units.  Placement 2: XCD k takes the k-th
// eighth of every chain's units.
void build_maps(qcd_batch* B, std::vector<uint2>& map)
{
    const uint2 idle = make_uint2(~0u, 0u);
    for (auto& bk : B->buckets) {
        std::vector<int> order = bk.members;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return B->chains[a]->nb_fused > B->chains[b]->nb_fused;
        });
        for (int pl = 0; pl < qcd_batch::kPlacements; ++pl) {
            std::vector<std::vector<uint2>> xcd(8);
            std::vector<uint64_t> load(8, 0);
            for (int c : order) {
                const uint32_t nu = B->chains[c]->nb_fused;
                if (pl == 2) {
                    for (uint32_t k = 0; k < 8; ++k)
                        for (uint32_t u = (uint32_t)((uint64_t)nu * k / 8); u < (uint32_t)((uint64_t)nu * (k + 1) / 8); ++u)
                            xcd[k].push_back(make_uint2((uint32_t)c, u));
                    continue;
                }
                const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                if (pl == 1) {
                    for (uint32_t u : cross_section_units(B->chains[c])) xcd[k].push_back(make_uint2((uint32_t)c, u));
                } else {
                    for (uint32_t u = 0; u < nu; ++u) xcd[k].push_back(make_uint2((uint32_t)c, u));
                }
                load[k] += nu;
            }
            size_t maxlen = 0;
            for (auto& l : xcd) maxlen = std::max(maxlen, l.size());
            bk.off3[pl] = (uint32_t)map.size();
            bk.len3[pl] = (uint32_t)(8 * maxlen);
            for (uint32_t b = 0; b < bk.len3[pl]; ++b) {
                const auto& l = xcd[b % 8];
                map.push_back(b / 8 < l.size() ? l[b / 8] : idle);
            }
            const size_t rows = (maxlen + kWaves1 - 1) / kWaves1;
            bk.off1[pl] = (uint32_t)map.size();
            bk.len1[pl] = (uint32_t)(8 * rows * kWaves1);
            for (uint32_t b = 0; b < 8 * rows; ++b)
                for (int w = 0; w < kWaves1; ++w) {
                    const auto& l = xcd[b % 8];
                    const size_t e = (b / 8) * kWaves1 + w;
                    map.push_back(e < l.size() ? l[e] : idle);
                }
        }
    }
    B->off_units = (uint32_t)map.size();
    for (size_t c = 0; c < B->chains.size(); ++c)
        for (uint32_t u = 0; u < B->chains[c]->nb_fused; ++u) map.push_back(make_uint2((uint32_t)c, u));
    B->len_units = (uint32_t)map.size() - B->off_units;
    B->off_poly = (uint32_t)map.size();
    for (size_t c = 0; c < B->chains.size(); ++c)
        for (uint32_t b = 0; b < B->chains[c]->poly_blocks; ++b) map.push_back(make_uint2((uint32_t)c, b));
    B->len_poly = (uint32_t)map.size() - B->off_poly;
}

// Rebuilds the merged maps from the buckets' current shapes and placements
// (host calls only: never inside a capture).
int build_merged(qcd_batch* B)
{
    std::vector<uint2> m;
    for (int kind = 0; kind < 2; ++kind)
        for (int si = 0; si < 2; ++si) {
            const int shape = si == 0 ? 3 : 1;
            B->moff[kind][si] = (uint32_t)m.size();
            for (const auto& bk : B->buckets) {
                if (bk.shape[kind] != shape) continue;
                const int pl = bk.place[kind];
                const uint32_t off = shape == 3 ? bk.off3[pl] : bk.off1[pl], len = shape == 3 ? bk.len3[pl] : bk.len1[pl];
                m.insert(m.end(), B->h_map.begin() + off, B->h_map.begin() + off + len);
            }
            B->mlen[kind][si] = (uint32_t)m.size() - B->moff[kind][si];
        }
    if (B->d_mmap) {   // kept until the batch is destroyed: a captured cycle may hold it
        B->mmap_old.push_back(B->d_mmap);
        B->d_mmap = nullptr;
    }
    if (m.empty()) return QCD_OK;
    if (hipMalloc((void**)&B->d_mmap, sizeof(uint2) * m.size()) != hipSuccess) {
        B->d_mmap = nullptr;
        return fail(QCD_E_NOMEM, "merged batch maps");
    }
    HIPCHK(hipMemcpy(B->d_mmap, m.data(), sizeof(uint2) * m.size(), hipMemcpyHostToDevice));
    return QCD_OK;
}

// One launch over `len` map entries of the given shape (entries per
// workgroup: 1 for shape 3, kWaves1 for shape 1).
template <typename R, int N, int NREAL, bool OVER, int RK>
void launch_batch_map(qcd_batch* B, int shape, const uint2* map, uint32_t len, int parity, int mu, int meas_mode,
                      const LaunchOpt& o)
{
    if (shape == 1) {
        const dim3 grid(len / kWaves1), block(kWaves1 * kLinks);
        if (meas_mode)
            launch(k_update_batch1<R, N, NREAL, OVER, RK, true>, grid, block, B->stream, o, (const ChainDesc*)B->d_desc,
                   map, parity, mu, meas_mode, B->rng_cur);
        else
            launch(k_update_batch1<R, N, NREAL, OVER, RK, false>, grid, block, B->stream, o, (const ChainDesc*)B->d_desc,
                   map, parity, mu, meas_mode, B->rng_cur);
    } else {
        const dim3 grid(len), block(kUpdThreads);
        if (meas_mode)
            launch(k_update_batch3<R, N, NREAL, OVER, RK, true>, grid, block, B->stream, o, (const ChainDesc*)B->d_desc,
                   map, parity, mu, meas_mode, B->rng_cur);
        else
            launch(k_update_batch3<R, N, NREAL, OVER, RK, false>, grid, block, B->stream, o, (const ChainDesc*)B->d_desc,
                   map, parity, mu, meas_mode, B->rng_cur);
    }
}

template <typename R, int N, int NREAL, bool OVER, int RK>
void launch_batch_merged(qcd_batch* B, int parity, int mu, int meas_mode, const LaunchOpt& o)
{
    const int kind = OVER ? 1 : 0;
    for (int si = 0; si < 2; ++si)
        if (B->mlen[kind][si])
            launch_batch_map<R, N, NREAL, OVER, RK>(B, si == 0 ? 3 : 1, B->d_mmap + B->moff[kind][si],
                                                    B->mlen[kind][si], parity, mu, meas_mode, o);
}

template <typename R, int N, int NREAL, bool OVER, int RK>
void launch_batch_update(qcd_batch* B, const qcd_batch::Bucket& bk, int parity, int mu, int meas_mode, const LaunchOpt& o)
{
    const int shape = bk.shape[OVER ? 1 : 0], pl = bk.place[OVER ? 1 : 0];
    if (shape == 1)
        launch_batch_map<R, N, NREAL, OVER, RK>(B, 1, B->d_map + bk.off1[pl], bk.len1[pl], parity, mu, meas_mode, o);
    else
        launch_batch_map<R, N, NREAL, OVER, RK>(B, 3, B->d_map + bk.off3[pl], bk.len3[pl], parity, mu, meas_mode, o);
}

// One (parity, mu) half-step of every chain; `only` >= 0 restricts it to one
// bucket (autotuning).
template <typename R, int N, int NREAL>
int batch_halfstep_rn(qcd_batch* B, int parity, int mu, int kind, int meas_mode, int only, const LaunchOpt& o)
{
    const qcd_params& p = B->p;
    if (kind == QCD_HEATBATH && p.prng == QCD_PRNG_RANLUX) {
        const uint32_t S = B->chains[0]->s_quads;
        const dim3 grid(B->len_units), block(kLinks);
        const uint2* map = B->d_map + B->off_units;
        switch (p.ranlux_luxury) {
            case 0: hipLaunchKernelGGL(k_ranlux_produce_batch<0>, grid, block, 0, B->stream, B->d_desc, map, S); break;
            case 1: hipLaunchKernelGGL(k_ranlux_produce_batch<1>, grid, block, 0, B->stream, B->d_desc, map, S); break;
            case 2: hipLaunchKernelGGL(k_ranlux_produce_batch<2>, grid, block, 0, B->stream, B->d_desc, map, S); break;
            case 3: hipLaunchKernelGGL(k_ranlux_produce_batch<3>, grid, block, 0, B->stream, B->d_desc, map, S); break;
            default: hipLaunchKernelGGL(k_ranlux_produce_batch<4>, grid, block, 0, B->stream, B->d_desc, map, S); break;
        }
    }
    if (only < 0 && B->merge) {
        if (kind == QCD_OVERRELAX) launch_batch_merged<R, N, NREAL, true, RK_NONE>(B, parity, mu, meas_mode, o);
        else if (p.prng == QCD_PRNG_RANLUX) launch_batch_merged<R, N, NREAL, false, RK_BUF>(B, parity, mu, meas_mode, o);
        else if (p.prng == QCD_PRNG_XOR128) launch_batch_merged<R, N, NREAL, false, RK_XOR128>(B, parity, mu, meas_mode, o);
        else launch_batch_merged<R, N, NREAL, false, RK_PM>(B, parity, mu, meas_mode, o);
    } else {
        for (size_t k = 0; k < B->buckets.size(); ++k) {
            if (only >= 0 && (int)k != only) continue;
            const auto& bk = B->buckets[k];
            if (kind == QCD_OVERRELAX) launch_batch_update<R, N, NREAL, true, RK_NONE>(B, bk, parity, mu, meas_mode, o);
            else if (p.prng == QCD_PRNG_RANLUX) launch_batch_update<R, N, NREAL, false, RK_BUF>(B, bk, parity, mu, meas_mode, o);
            else if (p.prng == QCD_PRNG_XOR128) launch_batch_update<R, N, NREAL, false, RK_XOR128>(B, bk, parity, mu, meas_mode, o);
            else launch_batch_update<R, N, NREAL, false, RK_PM>(B, bk, parity, mu, meas_mode, o);
        }
    }
    HIPCHK(hipGetLastError());
    if (kind == QCD_HEATBATH && p.prng != QCD_PRNG_RANLUX) {
        B->rng_cur ^= 1;   // every chain's states went to its other buffer
        for (auto* L : B->chains) L->rng_cur = B->rng_cur;
    }
    return QCD_OK;
}

// The batch kernels take one stream-buffer index for every chain: a chain
// swept on its own through its lattice handle would have flipped only its
// own, so the batch refuses to continue from such a state.
int batch_streams_in_step(const qcd_batch* B)
{
    for (const auto* L : B->chains)
        if (L->rng_cur != B->rng_cur || L->s_quads != B->chains[0]->s_quads)
            return fail(QCD_E_STATE, "a batched chain was stepped or re-sized on its own");
    return QCD_OK;
}

int batch_sweep(qcd_batch* B, int kind, bool fuse_plaquette)
{
    if (const int rc = batch_streams_in_step(B)) return rc;
    for (int parity = 1; parity >= 0; --parity)
        for (int mu = 0; mu < 4; ++mu) {
            const int meas = (fuse_plaquette && parity == 0 && mu >= 1) ? mu : 0;
            const int rc = DISPATCH(B, batch_halfstep_rn, B, parity, mu, kind, meas, -1, LaunchOpt{});
            if (rc) return rc;
        }
    return QCD_OK;
}

template <typename R, int N, int NREAL>
int batch_polyakov_rn(qcd_batch* B)
{
    hipLaunchKernelGGL((k_polyakov_batch<R, N, NREAL>), dim3(B->len_poly), dim3(kBlock), 0, B->stream,
                       (const ChainDesc*)B->d_desc, (const uint2*)(B->d_map + B->off_poly));
    HIPCHK(hipGetLastError());
    return QCD_OK;
}

// n_hb heatbath + n_or overrelaxation sweeps of every chain, then the
// plaquette (fused into the last sweep) and Polyakov records of every chain.
int batch_enqueue_cycle(qcd_batch* B, int n_hb, int n_or, int mask)
{
    if (const int rc = batch_streams_in_step(B)) return rc;
    const int total = n_hb + n_or;
    const bool fuse = (mask & QCD_MEAS_PLAQUETTE) && total > 0;
    for (int k = 0; k < total; ++k) {
        const int rc = batch_sweep(B, k < n_hb ? QCD_HEATBATH : QCD_OVERRELAX, fuse && k == total - 1);
        if (rc) return rc;
    }
    int flags = fuse ? 1 : 0;
    if (mask & QCD_MEAS_POLYAKOV) {
        const int rc = DISPATCH(B, batch_polyakov_rn, B);
        if (rc) return rc;
        flags |= 4;
    }
    if (flags) {
        hipLaunchKernelGGL(k_finalize_batch, dim3((uint32_t)B->chains.size()), dim3(kBlock), 0, B->stream,
                           (const ChainDesc*)B->d_desc, flags);
        HIPCHK(hipGetLastError());
    }
    return QCD_OK;
}

}  // namespace

extern "C" {

int qcd_batch_create(const qcd_params* common, int nchains, const qcd_chain_desc* chains, qcd_batch** out)
{
    if (!common || !chains || !out || nchains < 1) return fail(QCD_E_ARG, "null argument or no chains");
    if (common->prng != QCD_PRNG_XOR128 && common->prng != QCD_PRNG_PM && common->prng != QCD_PRNG_RANLUX)
        return fail(QCD_E_UNSUPPORTED, "batched chains need per-site streams (XOR128, PM or RANLUX)");
    if (common->prng_instances)
        return fail(QCD_E_UNSUPPORTED, "batched chains run PRNGSTEP generator instances (prng_instances 0)");
    *out = nullptr;
    qcd_batch* B = new qcd_batch;
    B->p = *common;
    for (int c = 0; c < nchains; ++c) {
        qcd_params q = *common;
        for (int k = 0; k < 4; ++k) q.n[k] = chains[c].n[k];
        q.beta = chains[c].beta;
        q.randseries = chains[c].randseries;
        qcd_lattice* L = nullptr;
        const int rc = qcd_lattice_create(&q, &L);
        if (rc) {
            qcd_batch_destroy(B);
            return rc;
        }
        B->chains.push_back(L);
        L->batch_owned = true;
        if (L->s_quads != B->chains[0]->s_quads) {   // one produce() size for the batched RANLUX launch
            qcd_batch_destroy(B);
            return fail(QCD_E_UNSUPPORTED, "batched chains need one stream production size");
        }
    }
    // buckets: chains of equal spatial extent
    for (int c = 0; c < nchains; ++c) {
        const qcd_lattice* L = B->chains[c];
        int found = -1;
        for (size_t k = 0; k < B->buckets.size(); ++k) {
            const qcd_lattice* M = B->chains[B->buckets[k].members[0]];
            if (M->g.n1 == L->g.n1 && M->g.n2 == L->g.n2 && M->g.n3 == L->g.n3) found = (int)k;
        }
        if (found < 0) {
            B->buckets.emplace_back();
            found = (int)B->buckets.size() - 1;
        }
        B->buckets[found].members.push_back(c);
    }
    std::vector<uint2> map;
    build_maps(B, map);
    std::vector<ChainDesc> desc(nchains);
    for (int c = 0; c < nchains; ++c) {
        const qcd_lattice* L = B->chains[c];
        desc[c].a = chain_base_args(L);
        desc[c].rng_buf[0] = L->rng_buf[0];
        desc[c].rng_buf[1] = L->rng_buf[1];
        desc[c].rl_state = L->rl_state;
        desc[c].nunits = L->nb_fused;
        desc[c].part_poly = L->part_poly;
        desc[c].fin = chain_fin_args(L);
    }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc((void**)&B->d_desc, sizeof(ChainDesc) * nchains)) != hipSuccess ||
        (e = hipMalloc((void**)&B->d_map, sizeof(uint2) * map.size())) != hipSuccess ||
        (e = hipMemcpy(B->d_desc, desc.data(), sizeof(ChainDesc) * nchains, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(B->d_map, map.data(), sizeof(uint2) * map.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        qcd_batch_destroy(B);
        return fail(QCD_E_HIP, "batch set-up: %s", hipGetErrorString(e));
    }
    B->h_map = std::move(map);
    if (int rc = build_merged(B)) {
        qcd_batch_destroy(B);
        return rc;
    }
    *out = B;
    return QCD_OK;
}

int qcd_batch_destroy(qcd_batch* B)
{
    if (!B) return QCD_OK;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    for (auto* L : B->chains) {
        L->batch_owned = false;
        qcd_lattice_destroy(L);
    }
    (void)hipFree(B->d_desc);
    (void)hipFree(B->d_map);
    (void)hipFree(B->d_mmap);
    for (uint2* m : B->mmap_old) (void)hipFree(m);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return QCD_OK;
}

int qcd_batch_sync(qcd_batch* B)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(B->stream));
    for (auto* L : B->chains) HIPCHK(hipStreamSynchronize(L->stream));
    return QCD_OK;
}

int qcd_batch_init(qcd_batch* B, int start)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(B->stream));
    for (auto* L : B->chains) {
        int rc = QCD_OK;
        if (start == QCD_START_HOT) {
            rc = qcd_prng_init(L);
            if (!rc) rc = qcd_lattice_init(L, start);
        } else if (start == QCD_START_COLD) {
            rc = qcd_lattice_init(L, start);
            if (!rc) rc = qcd_prng_init(L);
        } else {
            return fail(QCD_E_UNSUPPORTED, "batched chains start hot or cold");
        }
        if (rc) return rc;
        if (L->rng_cur != 0) return fail(QCD_E_STATE, "internal: stream buffer parity");
    }
    B->rng_cur = 0;
    for (auto* L : B->chains) HIPCHK(hipStreamSynchronize(L->stream));
    return QCD_OK;
}

int qcd_batch_sweep(qcd_batch* B, int kind)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (kind != QCD_HEATBATH && kind != QCD_OVERRELAX) return fail(QCD_E_ARG, "bad update kind");
    return batch_sweep(B, kind, false);
}

int qcd_batch_cycle_run(qcd_batch* B, int n_hb, int n_or, int mask)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (n_hb < 0 || n_or < 0 || n_hb + n_or < 1) return fail(QCD_E_ARG, "a cycle needs at least one sweep");
    return batch_enqueue_cycle(B, n_hb, n_or, mask);
}

int qcd_batch_cycle_capture(qcd_batch* B, int n_hb, int n_or, int mask)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (n_hb < 0 || n_or < 0 || n_hb + n_or < 1) return fail(QCD_E_ARG, "a cycle needs at least one sweep");
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        B->exec = nullptr;
    }
    if (B->graph) {
        (void)hipGraphDestroy(B->graph);
        B->graph = nullptr;
    }
    HIPCHK(hipStreamSynchronize(B->stream));
    HIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue_cycle(B, n_hb, n_or, mask);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(QCD_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    B->graph = graph;
    HIPCHK(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0));
    return QCD_OK;
}

int qcd_batch_cycle_launch(qcd_batch* B, int repeats)
{
    if (!B || !B->exec) return fail(QCD_E_STATE, "no captured cycle");
    for (int r = 0; r < repeats; ++r) HIPCHK(hipGraphLaunch(B->exec, B->stream));
    return QCD_OK;
}

int qcd_batch_cycle_time(qcd_batch* B, int repeats, double* ms)
{
    if (!B || !B->exec) return fail(QCD_E_STATE, "no captured cycle");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(B->stream));
    HIPCHK(hipEventRecord(a, B->stream));
    for (int r = 0; r < repeats; ++r) HIPCHK(hipGraphLaunch(B->exec, B->stream));
    HIPCHK(hipEventRecord(b, B->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = f;
    return QCD_OK;
}

// Times each bucket's half-steps with each kernel shape and keeps the faster
// shape per bucket and update kind.  Link tables and generator states are
// saved before and restored after, so tuning does not change any chain.
int qcd_batch_autotune(qcd_batch* B, int repeats, int* shapes_out)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (repeats < 1) return fail(QCD_E_ARG, "repeats must be >= 1");
    HIPCHK(hipStreamSynchronize(B->stream));
    // save
    std::vector<void*> saved;
    auto save = [&](const void* src, size_t bytes) -> bool {
        void* d = nullptr;
        if (!src || !bytes) {
            saved.push_back(nullptr);
            return true;
        }
        if (hipMalloc(&d, bytes) != hipSuccess) return false;
        saved.push_back(d);
        return hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, B->stream) == hipSuccess;
    };
    bool ok = true;
    for (auto* L : B->chains) {
        ok = ok && save(L->table, L->table_bytes);
        ok = ok && save(L->rng_buf[0], sizeof(uint4) * L->n_streams);
        ok = ok && save(L->rng_buf[1], sizeof(uint4) * L->n_streams);
        ok = ok && save(L->rl_state, L->rl_state ? sizeof(float4) * 7 * L->n_streams : 0);
    }
    const int rng_cur0 = B->rng_cur;
    hipEvent_t e0, e1;
    int rc = QCD_OK;
    if (ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        for (size_t k = 0; k < B->buckets.size() && !rc; ++k)
            for (int kind = 0; kind < 2 && !rc; ++kind) {
                float best = 1e30f;
                int best_shape = 3, best_place = 0;
                for (int cand = 0; cand < 2 * qcd_batch::kPlacements && !rc; ++cand) {
                    const int shape = (cand % 2) ? 1 : 3, place = cand / 2;
                    B->buckets[k].shape[kind] = shape;
                    B->buckets[k].place[kind] = place;
                    // one untimed launch, then `repeats` timed ones
                    rc = DISPATCH(B, batch_halfstep_rn, B, 1, 0, kind == 0 ? QCD_HEATBATH : QCD_OVERRELAX, 0, (int)k,
                                  LaunchOpt{});
                    (void)hipEventRecord(e0, B->stream);
                    for (int r = 0; r < repeats && !rc; ++r)
                        rc = DISPATCH(B, batch_halfstep_rn, B, r & 1, (r >> 1) & 3,
                                      kind == 0 ? QCD_HEATBATH : QCD_OVERRELAX, 0, (int)k, LaunchOpt{});
                    (void)hipEventRecord(e1, B->stream);
                    (void)hipEventSynchronize(e1);
                    float t = 0.f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    if (t < best) {
                        best = t;
                        best_shape = shape;
                        best_place = place;
                    }
                }
                B->buckets[k].shape[kind] = best_shape;
                B->buckets[k].place[kind] = best_place;
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        rc = fail(QCD_E_NOMEM, "autotune: no room to save the chains");
    }
    // restore
    size_t s = 0;
    for (auto* L : B->chains) {
        const void* dst[4] = {L->table, L->rng_buf[0], L->rng_buf[1], L->rl_state};
        const size_t bytes[4] = {L->table_bytes, sizeof(uint4) * L->n_streams, sizeof(uint4) * L->n_streams,
                                 L->rl_state ? sizeof(float4) * 7 * L->n_streams : 0};
        for (int j = 0; j < 4; ++j, ++s)
            if (s < saved.size() && saved[s] && bytes[j])
                (void)hipMemcpyAsync((void*)dst[j], saved[s], bytes[j], hipMemcpyDeviceToDevice, B->stream);
    }
    (void)hipStreamSynchronize(B->stream);
    for (void* d : saved) (void)hipFree(d);
    B->rng_cur = rng_cur0;
    for (auto* L : B->chains) L->rng_cur = rng_cur0;
    if (rc) return rc;
    if (shapes_out)
        for (size_t k = 0; k < B->buckets.size(); ++k) {
            shapes_out[2 * k] = B->buckets[k].shape[0];
            shapes_out[2 * k + 1] = B->buckets[k].shape[1];
        }
    // a captured cycle keeps the shapes it was captured with
    return build_merged(B);
}

int qcd_batch_set_tuning(qcd_batch* B, int bucket, int shape_hb, int shape_or, int place_hb, int place_or)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (bucket < 0 || bucket >= (int)B->buckets.size()) return fail(QCD_E_ARG, "bucket %d out of range", bucket);
    for (int v : {shape_hb, shape_or})
        if (v != 1 && v != 3) return fail(QCD_E_ARG, "kernel shape %d: 1 or 3", v);
    for (int v : {place_hb, place_or})
        if (v < 0 || v >= qcd_batch::kPlacements) return fail(QCD_E_ARG, "placement %d: 0..%d", v, qcd_batch::kPlacements - 1);
    auto& bk = B->buckets[bucket];
    bk.shape[0] = shape_hb;
    bk.shape[1] = shape_or;
    bk.place[0] = place_hb;
    bk.place[1] = place_or;
    return build_merged(B);   // a captured cycle keeps the choices it was captured with
}

int qcd_batch_set_merge(qcd_batch* B, int on)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    B->merge = on != 0;
    return QCD_OK;   // a captured cycle keeps the launches it was captured with
}

int qcd_batch_get_tuning(qcd_batch* B, int bucket, int* out)
{
    if (!B || !out) return fail(QCD_E_ARG, "null argument");
    if (bucket < 0 || bucket >= (int)B->buckets.size()) return fail(QCD_E_ARG, "bucket %d out of range", bucket);
    const auto& bk = B->buckets[bucket];
    out[0] = bk.shape[0];
    out[1] = bk.shape[1];
    out[2] = bk.place[0];
    out[3] = bk.place[1];
    const qcd_lattice* L = B->chains[bk.members[0]];
    out[4] = L->p.n[0];
    out[5] = (int)bk.members.size();
    return QCD_OK;
}

int qcd_batch_info(qcd_batch* B, int* nchains, int* nbuckets)
{
    if (!B) return fail(QCD_E_ARG, "null batch");
    if (nchains) *nchains = (int)B->chains.size();
    if (nbuckets) *nbuckets = (int)B->buckets.size();
    return QCD_OK;
}

int qcd_batch_chain(qcd_batch* B, int index, qcd_lattice** lat)
{
    if (!B || !lat) return fail(QCD_E_ARG, "null argument");
    if (index < 0 || index >= (int)B->chains.size()) return fail(QCD_E_ARG, "chain %d out of range", index);
    HIPCHK(hipStreamSynchronize(B->stream));
    *lat = B->chains[index];
    return QCD_OK;
}

}  // extern "C"
