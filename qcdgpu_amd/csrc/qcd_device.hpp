// qcd_device.hpp -- device-side building blocks for the SU(N) update and
// measurement kernels on gfx950: complex/SU(N) algebra in registers, the
// checkerboard-compacted link layout, and the per-stream generators.
//
// Link layout in HBM (one table per lattice):
//   plane-major, checkerboard-compacted structure of arrays.  For parity p,
//   direction mu and complex component k (row-major entry of the N x N
//   matrix; 9 for SU(3) 18-real, 6 for SU(3) 12-real, 2 for SU(2)), the
//   values of all Vh = V/2 sites of that parity are contiguous:
//       table[((p * 4 + mu) * NPLANE + k) * vh_pad + i]      (float2/double2)
//   i is the compact index of site s: i = s >> 1 (y is the fastest
//   coordinate and N2 is even, so each pair (2i, 2i+1) holds one site of each
//   parity).  One wavefront reads 64 consecutive complex values per plane:
//   512 B (FP32) or 1 KiB (FP64) fully coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qcd {

// ------------------------------------------------------------------ complex
template <typename R> struct C {
    R re, im;
};
template <typename R> __device__ __forceinline__ C<R> cmk(R a, R b) { return C<R>{a, b}; }
template <typename R> __device__ __forceinline__ C<R> cadd(C<R> a, C<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ C<R> csub(C<R> a, C<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __device__ __forceinline__ C<R> cmul(C<R> a, C<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename R> __device__ __forceinline__ C<R> cmulc(C<R> a, C<R> b)
{
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ C<R> ccmul(C<R> a, C<R> b)
{
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <typename R> __device__ __forceinline__ C<R> cconj(C<R> a) { return {a.re, -a.im}; }

// ------------------------------------------------------------------ matrices
template <typename R, int N> struct M {
    C<R> e[N][N];
};

template <typename R, int N> __device__ __forceinline__ void m_zero(M<R, N>& a)
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) a.e[i][j] = {R(0), R(0)};
}

template <typename R, int N> __device__ __forceinline__ void m_unit(M<R, N>& a)
{
    m_zero(a);
#pragma unroll
    for (int i = 0; i < N; ++i) a.e[i][i].re = R(1);
}

// Complex multiply-accumulate as four FMAs: s += a * b, s += a * conj(b),
// s += conj(a) * b.  The matrix products below accumulate entry by entry
// with these, so a 3x3 product is 27 x 4 FMAs and no separate adds.
template <typename R> __device__ __forceinline__ void cfma(C<R>& s, C<R> a, C<R> b)
{
    s.re = fma(a.re, b.re, s.re);
    s.re = fma(-a.im, b.im, s.re);
    s.im = fma(a.re, b.im, s.im);
    s.im = fma(a.im, b.re, s.im);
}
template <typename R> __device__ __forceinline__ void cfmac(C<R>& s, C<R> a, C<R> b)
{
    s.re = fma(a.re, b.re, s.re);
    s.re = fma(a.im, b.im, s.re);
    s.im = fma(a.im, b.re, s.im);
    s.im = fma(-a.re, b.im, s.im);
}
template <typename R> __device__ __forceinline__ void ccfma(C<R>& s, C<R> a, C<R> b)
{
    s.re = fma(a.re, b.re, s.re);
    s.re = fma(a.im, b.im, s.re);
    s.im = fma(a.re, b.im, s.im);
    s.im = fma(-a.im, b.re, s.im);
}

// c = a * b
template <typename R, int N> __device__ __forceinline__ M<R, N> m_mul(const M<R, N>& a, const M<R, N>& b)
{
    M<R, N> c;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            C<R> s = cmul(a.e[i][0], b.e[0][j]);
#pragma unroll
            for (int k = 1; k < N; ++k) cfma(s, a.e[i][k], b.e[k][j]);
            c.e[i][j] = s;
        }
    return c;
}

// c = a * b^dagger
template <typename R, int N> __device__ __forceinline__ M<R, N> m_mul_nd(const M<R, N>& a, const M<R, N>& b)
{
    M<R, N> c;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            C<R> s = cmulc(a.e[i][0], b.e[j][0]);
#pragma unroll
            for (int k = 1; k < N; ++k) cfmac(s, a.e[i][k], b.e[j][k]);
            c.e[i][j] = s;
        }
    return c;
}

// c = a^dagger * b
template <typename R, int N> __device__ __forceinline__ M<R, N> m_mul_dn(const M<R, N>& a, const M<R, N>& b)
{
    M<R, N> c;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            C<R> s = ccmul(a.e[0][i], b.e[0][j]);
#pragma unroll
            for (int k = 1; k < N; ++k) ccfma(s, a.e[k][i], b.e[k][j]);
            c.e[i][j] = s;
        }
    return c;
}

// acc += a * b^dagger
template <typename R, int N> __device__ __forceinline__ void m_acc_nd(M<R, N>& acc, const M<R, N>& a, const M<R, N>& b)
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
            for (int k = 0; k < N; ++k) cfmac(acc.e[i][j], a.e[i][k], b.e[j][k]);
}

// acc += a^dagger * b
template <typename R, int N> __device__ __forceinline__ void m_acc_dn(M<R, N>& acc, const M<R, N>& a, const M<R, N>& b)
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
            for (int k = 0; k < N; ++k) ccfma(acc.e[i][j], a.e[k][i], b.e[k][j]);
}

// Staple terms of SU(3) links with two computed rows: a product of SU(3)
// matrices is in SU(3), so its third row is conj(row 1 x row 2).
// F = f1 f2^+ f3^+: rows 0, 1 of f1 f2^+, then rows 0, 1 of that times f3^+.
template <typename R>
__device__ __forceinline__ M<R, 3> su3_fwd_term(const M<R, 3>& f1, const M<R, 3>& f2, const M<R, 3>& f3)
{
    M<R, 3> t, F;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            C<R> s = cmulc(f1.e[i][0], f2.e[j][0]);
            cfmac(s, f1.e[i][1], f2.e[j][1]);
            cfmac(s, f1.e[i][2], f2.e[j][2]);
            t.e[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            C<R> s = cmulc(t.e[i][0], f3.e[j][0]);
            cfmac(s, t.e[i][1], f3.e[j][1]);
            cfmac(s, t.e[i][2], f3.e[j][2]);
            F.e[i][j] = s;
        }
    su3_row3(F);
    return F;
}
// B = (b1 b2)^+ b3: columns 0, 1 of b1 b2 are rows 0, 1 of (b1 b2)^+.
template <typename R>
__device__ __forceinline__ M<R, 3> su3_bwd_term(const M<R, 3>& b1, const M<R, 3>& b2, const M<R, 3>& b3)
{
    M<R, 3> t, B;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            C<R> s = cmul(b1.e[k][0], b2.e[0][c]);
            cfma(s, b1.e[k][1], b2.e[1][c]);
            cfma(s, b1.e[k][2], b2.e[2][c]);
            t.e[k][c] = s;
        }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            C<R> s = ccmul(t.e[0][r], b3.e[0][c]);
            ccfma(s, t.e[1][r], b3.e[1][c]);
            ccfma(s, t.e[2][r], b3.e[2][c]);
            B.e[r][c] = s;
        }
    su3_row3(B);
    return B;
}

// Re Tr(a * b) accumulated in double.
template <typename R, int N> __device__ __forceinline__ double m_retr_mul(const M<R, N>& a, const M<R, N>& b)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            s = fma((double)a.e[i][k].re, (double)b.e[k][i].re, s);
            s = fma(-(double)a.e[i][k].im, (double)b.e[k][i].im, s);
        }
    return s;
}

template <typename R, int N> __device__ __forceinline__ void m_acc(M<R, N>& a, const M<R, N>& b)
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) a.e[i][j] = cadd(a.e[i][j], b.e[i][j]);
}

// Re Tr(a * b^dagger) = sum_ij Re(a_ij conj(b_ij)), accumulated in double.
template <typename R, int N> __device__ __forceinline__ double m_retr_nd(const M<R, N>& a, const M<R, N>& b)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
            s += (double)a.e[i][j].re * (double)b.e[i][j].re + (double)a.e[i][j].im * (double)b.e[i][j].im;
    return s;
}

template <typename Rout, typename Rin, int N> __device__ __forceinline__ M<Rout, N> m_cast(const M<Rin, N>& a)
{
    M<Rout, N> b;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) b.e[i][j] = {(Rout)a.e[i][j].re, (Rout)a.e[i][j].im};
    return b;
}

// Row 3 of an SU(3) matrix from rows 1 and 2: conj(row1 x row2).
template <typename R> __device__ __forceinline__ void su3_row3(M<R, 3>& a)
{
    a.e[2][0] = cconj(csub(cmul(a.e[0][1], a.e[1][2]), cmul(a.e[0][2], a.e[1][1])));
    a.e[2][1] = cconj(csub(cmul(a.e[0][2], a.e[1][0]), cmul(a.e[0][0], a.e[1][2])));
    a.e[2][2] = cconj(csub(cmul(a.e[0][0], a.e[1][1]), cmul(a.e[0][1], a.e[1][0])));
}

// Gram-Schmidt reunitarisation in FP64 (rows 1-2 orthonormalised, row 3 =
// conj(row1 x row2)), the algorithm of suncpp/su3/algebra_su3.cpp:185-226.
// ------------------------------------------------------------------ fast math
// 1/sqrt(x), x > 0, in double: the hardware estimate refined by two Newton
// steps y <- y + (y/2)(1 - x y^2) (quadratic convergence; about one ulp).
__device__ __forceinline__ double rsqrt_d(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return y;
}

// sin and cos of 2*pi*x for |x| <= 1 (x in revolutions).  Quadrant
// n = rint(4x); f = x - n/4 lies in [-1/8, 1/8] and is exact; theta = 2*pi*f;
// Taylor polynomials through theta^9 (sin) and theta^8 (cos) have truncation
// error below 3e-8 on |theta| <= pi/4; the result is rotated by n quarter
// turns.
__host__ __device__ __forceinline__ void sincos2pi_f(float x, float* sn, float* cs)
{
    const float n = rintf(4.0f * x);
    const float f = fmaf(n, -0.25f, x);
    const float t = f * 6.28318530717958647692f;
    const float t2 = t * t;
    float s = fmaf(t2, 2.75573192e-6f, -1.98412698e-4f);
    s = fmaf(t2, s, 8.33333333e-3f);
    s = fmaf(t2, s, -1.66666667e-1f);
    s = fmaf(t2 * t, s, t);
    float c = fmaf(t2, 2.48015873e-5f, -1.38888889e-3f);
    c = fmaf(t2, c, 4.16666667e-2f);
    c = fmaf(t2, c, -0.5f);
    c = fmaf(t2, c, 1.0f);
    const int q = (int)n & 3;
    float so = s, co = c;
    if (q == 1) {
        so = c;
        co = -s;
    } else if (q == 2) {
        so = -s;
        co = -c;
    } else if (q == 3) {
        so = -c;
        co = s;
    }
    *sn = so;
    *cs = co;
}

__device__ __forceinline__ void su3_gram_schmidt(M<double, 3>& a)
{
    double nu = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) nu += a.e[0][k].re * a.e[0][k].re + a.e[0][k].im * a.e[0][k].im;
    double inu = rsqrt_d(nu);
#pragma unroll
    for (int k = 0; k < 3; ++k) a.e[0][k] = {a.e[0][k].re * inu, a.e[0][k].im * inu};
    C<double> sp = cmulc(a.e[1][0], a.e[0][0]);
#pragma unroll
    for (int k = 1; k < 3; ++k) sp = cadd(sp, cmulc(a.e[1][k], a.e[0][k]));
#pragma unroll
    for (int k = 0; k < 3; ++k) a.e[1][k] = csub(a.e[1][k], cmul(a.e[0][k], sp));
    double nv = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) nv += a.e[1][k].re * a.e[1][k].re + a.e[1][k].im * a.e[1][k].im;
    double inv = rsqrt_d(nv);
#pragma unroll
    for (int k = 0; k < 3; ++k) a.e[1][k] = {a.e[1][k].re * inv, a.e[1][k].im * inv};
    su3_row3(a);
}

// SU(2) normalisation of the quaternion (Re u11, Im u11, Re u12, Im u12).
__device__ __forceinline__ void su2_normalize(M<double, 2>& a)
{
    double n = a.e[0][0].re * a.e[0][0].re + a.e[0][0].im * a.e[0][0].im + a.e[0][1].re * a.e[0][1].re +
               a.e[0][1].im * a.e[0][1].im;
    double in = rsqrt_d(n);
    a.e[0][0] = {a.e[0][0].re * in, a.e[0][0].im * in};
    a.e[0][1] = {a.e[0][1].re * in, a.e[0][1].im * in};
    a.e[1][0] = {-a.e[0][1].re, a.e[0][1].im};
    a.e[1][1] = {a.e[0][0].re, -a.e[0][0].im};
}

// ------------------------------------------------------------------ layout
template <typename R> struct Vec2;
template <> struct Vec2<float> {
    using T = float2;
};
template <> struct Vec2<double> {
    using T = double2;
};

// Number of complex planes stored per link.
template <int N, int NREAL> struct Planes {
    static constexpr int value = NREAL / 2;
};

// Unsigned division by a run-time constant d >= 1 as multiply-high, shift
// and add (Granlund and Montgomery, PLDI 1994, round-up method with the
// 33-bit magic number folded into the add): with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1 and t = mulhi(m, n),
// n / d = (t + ((n - t) >> s1)) >> s2, s1 = 1, s2 = l - 1 (d = 1: m = 0,
// s1 = s2 = 0).  Exact for every 32-bit n.
struct FastDiv {
    uint32_t d, m, s1, s2;
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t t = __umulhi(m, n);
#else
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
#endif
        return (t + ((n - t) >> s1)) >> s2;
    }
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{d, 0u, 0u, 0u};
    if (d > 1) {
        uint32_t l = 0;
        while ((1ull << l) < d) ++l;
        f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
        f.s1 = 1;
        f.s2 = l - 1;
    }
    return f;
}

struct Geom {
    int n1, n2, n3, n4;   // extents along x, y, z, t
    int h;                // N2 / 2
    uint32_t vh;          // V / 2
    uint32_t vh_pad;      // plane stride (multiple of 64)
    FastDiv dh, d3, d4;   // division by h, N3, N4
};

// Twisted boundary: the Y link of every site with x = N1 - 1 (compact index
// >= from in either parity, x being the slowest coordinate) is read as
// diag(d) U.  Passed to the kernels after all other arguments, so that the
// argument layout of the untwisted kernels is unchanged.
struct Twist {
    int on = 0;
    uint32_t from = 0xffffffffu;
    double d[6] = {1.0, 0.0, 1.0, 0.0, 1.0, 0.0};   // (re, im) of the diagonal entries
};

struct Site {
    int x, y, z, t;
};

// Coordinates of the site of parity p with compact index i (y fastest).
__host__ __device__ __forceinline__ Site site_of(const Geom& g, uint32_t i, int p)
{
    Site s;
    const uint32_t r = g.dh.div(i);
    const uint32_t yh = i - r * (uint32_t)g.h;
    const uint32_t r2 = g.d3.div(r);
    s.z = (int)(r - r2 * (uint32_t)g.n3);
    const uint32_t r3 = g.d4.div(r2);
    s.t = (int)(r2 - r3 * (uint32_t)g.n4);
    s.x = (int)r3;
    s.y = 2 * (int)yh + ((s.x + s.z + s.t + p) & 1);
    return s;
}

// Neighbour steps along a wavefront-uniform direction d (so every select on
// d is scalar).  c = the site's coordinate along d, o = y & 1 of the site,
// i = its compact index.  x, z and t steps move the compact index by the
// direction's stride (wrapping at the boundary); a y step keeps or moves the
// half-row index depending on o.  A step along one direction leaves the
// coordinate along any other direction unchanged, and changes o only when
// it is a y step, so two-step neighbours chain two calls with the original
// c and o (the second direction is never y when the first one is).
__host__ __device__ __forceinline__ int coord(const Site& s, int d)
{
    return d == 0 ? s.x : (d == 1 ? s.y : (d == 2 ? s.z : s.t));
}
__host__ __device__ __forceinline__ uint32_t cstride(const Geom& g, int d)
{
    return d == 0 ? (uint32_t)g.h * (uint32_t)g.n3 * (uint32_t)g.n4
                  : (d == 2 ? (uint32_t)g.h : (uint32_t)g.h * (uint32_t)g.n3);
}
__host__ __device__ __forceinline__ int extent(const Geom& g, int d)
{
    return d == 0 ? g.n1 : (d == 1 ? g.n2 : (d == 2 ? g.n3 : g.n4));
}
__host__ __device__ __forceinline__ uint32_t step_fwd(const Geom& g, uint32_t i, int d, int c, uint32_t o)
{
    const int n = extent(g, d);
    if (d == 1) return (c < n - 1) ? i + o : i - (uint32_t)(g.h - 1);
    const uint32_t st = cstride(g, d);
    return (c < n - 1) ? i + st : i - st * (uint32_t)(n - 1);
}
__host__ __device__ __forceinline__ uint32_t step_bwd(const Geom& g, uint32_t i, int d, int c, uint32_t o)
{
    const int n = extent(g, d);
    if (d == 1) return (c > 0) ? i - (1u - o) : i + (uint32_t)(g.h - 1);
    const uint32_t st = cstride(g, d);
    return (c > 0) ? i - st : i + st * (uint32_t)(n - 1);
}

__host__ __device__ __forceinline__ uint32_t cindex(const Geom& g, int x, int y, int z, int t)
{
    return (uint32_t)(y >> 1) + (uint32_t)g.h * ((uint32_t)z + (uint32_t)g.n3 * ((uint32_t)t + (uint32_t)g.n4 * (uint32_t)x));
}

__host__ __device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Compact index of s + sign * e_dir, and of s + s1 * e_d1 + s2 * e_d2.
// Directions are selected with branches (no dynamically indexed arrays, so
// nothing spills to scratch when dir is a runtime value).
__host__ __device__ __forceinline__ void shift(const Geom& g, Site& s, int dir, int sign)
{
    if (dir == 0) s.x = wrap(s.x + sign, g.n1);
    else if (dir == 1) s.y = wrap(s.y + sign, g.n2);
    else if (dir == 2) s.z = wrap(s.z + sign, g.n3);
    else s.t = wrap(s.t + sign, g.n4);
}

__host__ __device__ __forceinline__ uint32_t nb_index(const Geom& g, Site s, int dir, int sign)
{
    shift(g, s, dir, sign);
    return cindex(g, s.x, s.y, s.z, s.t);
}

__host__ __device__ __forceinline__ uint32_t nb2_index(const Geom& g, Site s, int d1, int s1, int d2, int s2)
{
    shift(g, s, d1, s1);
    shift(g, s, d2, s2);
    return cindex(g, s.x, s.y, s.z, s.t);
}

template <typename R, int N, int NREAL> struct Table {
    using V2 = typename Vec2<R>::T;
    static constexpr int NP = Planes<N, NREAL>::value;
    V2* __restrict__ base;
    uint32_t vh_pad;

    // Plane k of (parity, mu): uniform across the wavefront whenever p and mu
    // are, so loads use a scalar base and a 32-bit per-lane byte offset.
    __device__ __forceinline__ const V2* plane(int p, int mu, int k) const
    {
        return base + (size_t)((p * 4 + mu) * NP + k) * vh_pad;
    }
    // The plane base goes through readfirstlane so the compiler keeps it in
    // SGPRs as an opaque value and selects the global_load saddr form
    // (scalar base + 32-bit lane offset) instead of a 64-bit add per load.
    using gchar = __attribute__((address_space(1))) char;
    __device__ __forceinline__ static gchar* sbase(const V2* pl)
    {
        const uint64_t b = (uint64_t)pl;
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
        return (gchar*)(((uint64_t)hi << 32) | lo);
    }
    __device__ __forceinline__ V2 ld(const V2* pl, uint32_t i) const
    {
        using gR = __attribute__((address_space(1))) const R;
        const uint32_t off = i * (uint32_t)sizeof(V2);
        const gR* q = (const gR*)(sbase(pl) + off);
        V2 v;
        v.x = q[0];
        v.y = q[1];
        return v;
    }

    // Generic load: p and mu may differ between the lanes of a wavefront.
    __device__ __forceinline__ V2 ld_any(const V2* pl, uint32_t i) const { return pl[i]; }

    template <bool UNIFORM>
    __device__ __forceinline__ M<R, N> load_t(int p, int mu, uint32_t i) const
    {
        M<R, N> a;
        if constexpr (N == 3) {
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                V2 v = UNIFORM ? ld(plane(p, mu, k), i) : ld_any(plane(p, mu, k), i);
                a.e[k / 3][k % 3] = {v.x, v.y};
            }
            if constexpr (NP == 6) su3_row3(a);
        } else {
            V2 v0 = UNIFORM ? ld(plane(p, mu, 0), i) : ld_any(plane(p, mu, 0), i);
            V2 v1 = UNIFORM ? ld(plane(p, mu, 1), i) : ld_any(plane(p, mu, 1), i);
            a.e[0][0] = {v0.x, v0.y};
            a.e[0][1] = {v1.x, v1.y};
            a.e[1][0] = {-v1.x, v1.y};
            a.e[1][1] = {v0.x, -v0.y};
        }
        return a;
    }
    // Stored planes only (12-real SU(3): rows 1 and 2; call finish() for
    // row 3), so that a batch of gathers is not split by arithmetic.
    // PRECONDITION: p and mu are the same in every active lane.
    __device__ __forceinline__ M<R, N> load_raw_u(int p, int mu, uint32_t i) const
    {
        M<R, N> a;
        if constexpr (N == 3) {
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                V2 v = ld(plane(p, mu, k), i);
                a.e[k / 3][k % 3] = {v.x, v.y};
            }
        } else {
            a = load_t<true>(p, mu, i);
        }
        return a;
    }
    __device__ __forceinline__ static void finish(M<R, N>& a)
    {
        if constexpr (N == 3 && NP == 6) su3_row3(a);
    }
    // Load with wavefront-uniform parity p and direction mu (scalar plane
    // base).  PRECONDITION: p and mu are the same in every active lane.
    __device__ __forceinline__ M<R, N> load_u(int p, int mu, uint32_t i) const { return load_t<true>(p, mu, i); }
    // Load where p or mu may vary across lanes.
    __device__ __forceinline__ M<R, N> load(int p, int mu, uint32_t i) const { return load_t<false>(p, mu, i); }

    // Store with wavefront-uniform p and mu (same precondition as load_u).
    __device__ __forceinline__ void store_u(int p, int mu, uint32_t i, const M<R, N>& a) const
    {
        const uint32_t off = i * (uint32_t)sizeof(V2);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            V2 v;
            v.x = a.e[k / N][k % N].re;
            v.y = a.e[k / N][k % N].im;
            using gRw = __attribute__((address_space(1))) R;
            gRw* q = (gRw*)(sbase(plane(p, mu, k)) + off);
            q[0] = v.x;
            q[1] = v.y;
        }
    }
};

// Twisted boundary conditions (lattice_table_3 / lattice_table_2 under TBC,
// su3_matrix_memory.cl:57-98, su2_matrix_memory.cl:56-88): the reference
// scales the stored rows of the Y link at x = N1 - 1 by e^{i(phi+omega)/2}
// and e^{i(omega-phi)/2} (SU(3); row 3, rebuilt from them, picks up
// e^{-i omega}), or multiplies by diag(e^{i phi/2}, e^{-i phi/2}) (SU(2)),
// every time the link is read; stores write the link as computed.  Here the
// whole matrix is scaled row by row after the load (the same matrix: for
// two-row storage the rebuilt third row of the scaled rows equals the
// third row scaled by e^{-i omega}).
template <typename R, int N>
__device__ __forceinline__ void twist_rows(M<R, N>& a, const Twist& t)
{
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const C<R> f{(R)t.d[2 * r], (R)t.d[2 * r + 1]};
#pragma unroll
        for (int c = 0; c < N; ++c) a.e[r][c] = cmul(f, a.e[r][c]);
    }
}
// The link of direction dir at compact index idx as every kernel reads it.
template <typename R, int N>
__device__ __forceinline__ void tbc(M<R, N>& a, const Twist& t, int dir, uint32_t idx)
{
    if (t.on && dir == 1 && idx >= t.from) twist_rows(a, t);
}

// ------------------------------------------------------------------ generators
// XOR128 (Marsaglia 2003), one step as prngcl_xor128.cl:62-74.
__host__ __device__ __forceinline__ uint32_t xor128_step(uint4& s)
{
    uint32_t t = s.x ^ (s.x << 11);
    s.x = s.y;
    s.y = s.z;
    s.z = s.w;
    s.w = (s.w ^ (s.w >> 19)) ^ (t ^ (t >> 8));
    return s.w;
}

__device__ __forceinline__ float4 xor128_quad(uint4& s)
{
    float4 q;
    q.x = (float)xor128_step(s) * 2.3283064365386962890625e-10f;   // / 2^32, exact scaling
    q.y = (float)xor128_step(s) * 2.3283064365386962890625e-10f;
    q.z = (float)xor128_step(s) * 2.3283064365386962890625e-10f;
    q.w = (float)xor128_step(s) * 2.3283064365386962890625e-10f;
    return q;
}

// Park-Miller on four independent lanes (prngcl_pm.cl:68-84); output is an
// exact scaling by 2^-31 (the reference divides by (float)2147483647.0).
// Jump-ahead by a fixed number of steps.  The xorshift128 step is linear
// over GF(2) in the 128 state bits (shifts and XORs only), so n steps are a
// fixed 128x128 bit matrix; cols[b] = the state reached from the state with
// only bit b set (bit b = bit b%32 of word b/32 of (x, y, z, w)), and the
// jumped state is the XOR of the columns of the set bits.  cols is
// wavefront-uniform (scalar loads).
// Words 2*half and 2*half+1 of the jumped state (half is wavefront-uniform,
// so two wavefronts can share one jump).
__device__ __forceinline__ uint2 xor128_jump_half(uint4 s, const uint4* cols_g, int half)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // Read through the constant address space so the uniform column loads
    // are scalar; o ^= (col & m) is one v_bitop3_b32 per word (truth table
    // 0x6c for operands (col, o, m)).
    using cu32 = __attribute__((address_space(4))) const uint32_t;
    cu32* cols = (cu32*)cols_g + 2 * half;
    uint32_t oa = 0u, ob = 0u;
    const uint32_t w[4] = {s.x, s.y, s.z, s.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll 8
        for (int b = 0; b < 32; ++b) {
            const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)w[q], b, 1);   // 0 or ~0
            cu32* c = cols + 4 * (q * 32 + b);
            oa = __builtin_amdgcn_bitop3_b32(c[0], oa, m, 0x6c);
            ob = __builtin_amdgcn_bitop3_b32(c[1], ob, m, 0x6c);
        }
    }
    return make_uint2(oa, ob);
#else
    (void)cols_g;
    return half ? make_uint2(s.z, s.w) : make_uint2(s.x, s.y);   // device-only; never called on the host
#endif
}

// The same jump four state bits at a time: tab[16 c + e] (c = 0..31, e =
// 0..15) is the XOR of the columns of bits 4c .. 4c+3 selected by e, so the
// jumped state is the XOR of 32 table entries, one per nibble of the state.
// tab is the table staged in LDS; words 2*half and 2*half+1 as above.
constexpr int kJumpCols = 128;   // uint4 columns
constexpr int kJumpTab = 512;    // uint4 table entries after the columns
__host__ __device__ __forceinline__ uint2 xor128_jump_tab_half(uint4 s, const uint4* tab, int half)
{
    const uint32_t w[4] = {s.x, s.y, s.z, s.w};
    const uint2* t = reinterpret_cast<const uint2*>(tab) + half;
    uint32_t oa = 0u, ob = 0u;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const uint32_t e = (w[c >> 3] >> (4 * (c & 7))) & 15u;
        const uint2 v = t[2 * (16 * c + (int)e)];
        oa ^= v.x;
        ob ^= v.y;
    }
    return make_uint2(oa, ob);
}

// Park-Miller (Schrage form) for 0 < s < m is s -> 16807 s mod m, so k steps
// are one multiplication by a_k = 16807^k mod m; the Schrage form maps 0 to m
// and keeps m, which the jump reproduces.
__host__ __device__ __forceinline__ uint32_t pm_mulmod(uint32_t s, uint32_t a)
{
    const uint64_t p = (uint64_t)s * a;                 // < 2^62
    uint64_t r = (p & 2147483647ull) + (p >> 31);        // < 2^32, == p mod m
    r = (r & 2147483647ull) + (r >> 31);
    return (r >= 2147483647ull) ? (uint32_t)(r - 2147483647ull) : (uint32_t)r;
}
__host__ __device__ __forceinline__ uint32_t pm_jump(uint32_t s, uint32_t ak)
{
    return (s == 0u || s == 2147483647u) ? 2147483647u : pm_mulmod(s, ak);
}

__host__ __device__ __forceinline__ uint32_t pm_step(uint32_t s)
{
    uint32_t t1 = 16807u * (s % 127773u);
    uint32_t t2 = 2836u * (s / 127773u);
    return (t1 > t2) ? t1 - t2 : t1 - t2 + 2147483647u;
}

__device__ __forceinline__ float4 pm_quad(uint4& s)
{
    s.x = pm_step(s.x);
    s.y = pm_step(s.y);
    s.z = pm_step(s.z);
    s.w = pm_step(s.w);
    const float k = 4.656612873077392578125e-10f;   // 2^-31
    return make_float4((float)s.x * k, (float)s.y * k, (float)s.z * k, (float)s.w * k);
}

// ------------------------------------------------------------------ RANLUX
// The reference's float subtract-with-borrow generator (prngcl_ranlux.cl)
// with the 24 lagged values in registers.  All streams of a lattice advance
// in lockstep (every produce call draws the same number of quads from every
// stream), so the position inside a 24-block is uniform and every ring index
// below is a compile-time constant after unrolling: quad k of a block uses
// ring positions i = 23 - 4k .. 20 - 4k with j = (i + 10) mod 24; a block
// ends at i = 23 again; the skip of P numbers is P / 24 full cycles, P mod 24
// single steps, then a renaming of the ring by P mod 24 positions so the
// next block starts at i = 23 (the RL_skip tails of prngcl_ranlux.cl:229-466).
constexpr float kRlTwom24 = 5.9604644775390625e-08f;
constexpr float kRlTwom12 = 2.44140625e-04f;

// The ring holds the 24-bit integers k of the values k 2^-24 (every RANLUX
// number is one, the seeds included), so a subtract-with-borrow step is one
// 32-bit subtraction with borrow in and out and a 24-bit mask: for
// s[J], s[I] < 2^24 and carry c in {0, 1}, s[J] - s[I] - c < 0 exactly when
// the 32-bit subtraction borrows, and then the masked result is
// s[J] - s[I] - c + 2^24 -- the float form's "uni += 1, carry = 2^-24"
// step for step, two VALU instructions (v_subb_co_u32 with the borrow in
// VCC, v_and_b32) against six for the float compare-and-select form.
// Outputs are converted to float (exact) and get the float form's
// small-value fix with the same float arithmetic.
template <int I, int J>
__host__ __device__ __forceinline__ uint32_t rl_swb(uint32_t (&s)[24], uint32_t& carry)
{
    uint32_t borrow;
    const uint32_t d = __builtin_subc(s[J], s[I], carry, &borrow) & 0xFFFFFFu;
    carry = borrow;
    s[I] = d;
    return d;
}

// One output: SWB at (I, J) then the small-value fix with the value at the
// decremented j pointer (prngcl_ranlux.cl:178-179).
template <int I>
__host__ __device__ __forceinline__ float rl_out(uint32_t (&s)[24], uint32_t& carry)
{
    constexpr int J = (I + 10) % 24;
    float uni = (float)rl_swb<I, J>(s, carry) * kRlTwom24;
    if (uni < kRlTwom12) {
        const float lo = (float)s[(J + 23) % 24] * kRlTwom24;
        uni = uni + kRlTwom24 * lo;
        if (uni == 0.0f) uni = kRlTwom24 * kRlTwom24;
    }
    return uni;
}

template <int K>
__host__ __device__ __forceinline__ float4 rl_quad(uint32_t (&s)[24], uint32_t& carry)
{
    float4 q;
    q.x = rl_out<23 - 4 * K>(s, carry);
    q.y = rl_out<22 - 4 * K>(s, carry);
    q.z = rl_out<21 - 4 * K>(s, carry);
    q.w = rl_out<20 - 4 * K>(s, carry);
    return q;
}

template <int T>
__host__ __device__ __forceinline__ void rl_steps(uint32_t (&s)[24], uint32_t& carry)
{
    if constexpr (T > 0) {
        constexpr int I = (23 - (T - 1) + 24) % 24;   // step T-1 of a cycle
        rl_steps<T - 1>(s, carry);
        (void)rl_swb<I, (I + 10) % 24>(s, carry);
    }
}

// Skip P = 24 * FULL + REM numbers from the start of a block, then rename.
template <int FULL, int REM>
__host__ __device__ __forceinline__ void rl_skip(uint32_t (&s)[24], uint32_t& carry)
{
    for (int f = 0; f < FULL; ++f) rl_steps<24>(s, carry);
    rl_steps<REM>(s, carry);
    if constexpr (REM != 0) {
        uint32_t t[24];
#pragma unroll
        for (int i = 0; i < 24; ++i) t[(i + REM) % 24] = s[i];
#pragma unroll
        for (int i = 0; i < 24; ++i) s[i] = t[i];
    }
}

// Luxury level -> skip: 24, 48, 97, 223, 389 (random.cpp:190-194).
template <int LUX> struct RlSkip;
template <> struct RlSkip<0> { static constexpr int full = 1, rem = 0; };
template <> struct RlSkip<1> { static constexpr int full = 2, rem = 0; };
template <> struct RlSkip<2> { static constexpr int full = 4, rem = 1; };
template <> struct RlSkip<3> { static constexpr int full = 9, rem = 7; };
template <> struct RlSkip<4> { static constexpr int full = 16, rem = 5; };

}  // namespace qcd
