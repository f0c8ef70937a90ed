/*
 * oracle.c -- CPU restatement of QCDGPU's SU(N) update and measurement path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Each function cites the reference file:line it restates
 * (paths relative to the reference's SU(N)/ tree).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   glibc rand sequence ........ golden fixture written with glibc 2.35
 *   RANLUX scalar stream ....... golden fixture from the reference's own
 *                                RL_produce_one_CPU (oracle/_ref)
 *   PM CPU twin ................ golden fixture from the reference's own
 *                                PM_produce_one_CPU + Park-Miller 1988 KAT
 *   XOR128 ..................... Marsaglia (2003) recurrence, independent
 *                                implementation in the test + published value
 *   update in GID / SEQ mode ... golden links from the reference's own suncpp
 *                                (lattice_update_{odd,even}) on 4^4 and 6x4x2x4
 *   plaquette .................. golden plqConf values + closed forms
 *   QUADS (GPU buffer) mode .... parity unpinned by the reference (its OpenCL
 *                                kernels cannot run here); shares all
 *                                arithmetic with SEQ mode and differs only in
 *                                the quad-fetch rule, which is restated from
 *                                su3_update_cl.cl:597-668
 *   overrelaxation ............. no reference code; pinned by invariants
 *                                (ReTr(U A) preserved, involution)
 *   Polyakov loop .............. closed forms (cold start, gauge invariance)
 *   F_munu ..................... closed form (constant abelian flux along
 *                                each generator: F = V sin b) and the
 *                                explicit Gell-Mann / Pauli matrix traces
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI2 6.2831853071795864769252867665590

/* ======================================================================== */
/* glibc seeding: the reference seeds everything through srand/rand          */
/* (random/random.cpp:235,465,508,655-658,677-680,868-876).                  */
/* ======================================================================== */
void orc_srand(unsigned seed) { srand(seed); }
int  orc_rand(void) { return rand(); }
void orc_rand_fill(unsigned seed, int count, int *out)
{
    srand(seed);
    for (int i = 0; i < count; ++i) out[i] = rand();
}

/* ======================================================================== */
/* XOR128 -- random/prngcl_xor128.cl:62-74 (step), :100-133 (kernel)         */
/* ======================================================================== */
uint32_t orc_xor128_step(uint32_t s[4])
{
    uint32_t t = s[0] ^ (s[0] << 11);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    return s[3];
}

/* random.cpp:227-241 (srand), :675-681 (XOR128_initialize_CPU: instance 0),
   :647-659 (other instances, 4 rand() each in x,y,z,w order).             */
void orc_xor128_seed_table(unsigned randseries, unsigned n, uint32_t *table)
{
    srand(randseries);
    for (unsigned g = 0; g < n; ++g)
        for (int c = 0; c < 4; ++c) table[4 * g + c] = (uint32_t)rand();
}

/* Kernel body: 4 steps per quad, w / 2^32 in float (prngcl_xor128.cl:120-129). */
void orc_xor128_produce(uint32_t *table, unsigned n, unsigned samples, float *quads)
{
    for (unsigned g = 0; g < n; ++g) {
        uint32_t *s = table + 4 * (size_t)g;
        for (unsigned i = 0; i < samples; ++i) {
            float *q = quads + 4 * ((size_t)i * n + g);
            for (int c = 0; c < 4; ++c) q[c] = (float)orc_xor128_step(s) / 4294967296.0f;
        }
    }
}

/* ======================================================================== */
/* Park-Miller -- random/prngcl_pm.cl:68-84 (step), :86-124 (kernel)         */
/* ======================================================================== */
uint32_t orc_pm_step(uint32_t s)
{
    uint32_t t1 = 16807u * (s % 127773u);
    uint32_t t2 = 2836u * (s / 127773u);
    return (t1 > t2) ? t1 - t2 : t1 - t2 + 2147483647u;
}

/* random.cpp:892-894 (CPU seed), :867-876 (table: instance 0 = CPU seed and
   3 more rand()%m, others 4 x rand()%m).                                   */
void orc_pm_seed_table(unsigned randseries, unsigned n, uint32_t *table)
{
    srand(randseries);
    for (unsigned g = 0; g < n; ++g)
        for (int c = 0; c < 4; ++c) table[4 * g + c] = (uint32_t)(rand() % 2147483647);
}

/* Kernel output is (float)seed / (float)2147483647.0; the divisor rounds to
   2^31, so the quotient is an exact power-of-two scaling.                  */
void orc_pm_produce(uint32_t *table, unsigned n, unsigned samples, float *quads)
{
    for (unsigned g = 0; g < n; ++g) {
        uint32_t *s = table + 4 * (size_t)g;
        for (unsigned i = 0; i < samples; ++i) {
            float *q = quads + 4 * ((size_t)i * n + g);
            for (int c = 0; c < 4; ++c) {
                s[c] = orc_pm_step(s[c]);
                q[c] = (float)s[c] / (float)2147483647.0;
            }
        }
    }
}

/* random.cpp:896-910: Schrage form in int, double division then float.     */
float orc_pm_one_cpu(int32_t *seed)
{
    uint32_t hi = (uint32_t)(*seed / 127773);
    uint32_t lo = (uint32_t)(*seed % 127773);
    int32_t test = (int32_t)(16807u * lo - 2836u * hi);
    *seed = (test > 0) ? test : test + 2147483647;
    return (float)((double)*seed / 2147483647.0);
}

/* ======================================================================== */
/* RANLUX -- scalar twin random.cpp:504-574; float4 kernel                   */
/* prngcl_ranlux.cl:138-157 (init), :160-201 (quad), :204-498 (skip), 534.   */
/* ======================================================================== */
#define RL_ICONS  2147483563
#define RL_TWO24  16777216
#define RL_TWOM24 0.000000059604644775390625f
#define RL_TWOM12 0.000244140625f

int orc_ranlux_nskip(int luxury)
{
    static const int p[5] = {24, 48, 97, 223, 389};   /* random.cpp:190-194 */
    return (luxury >= 0 && luxury < 5) ? p[luxury] : 24;
}

/* L'Ecuyer MLCG 40014/53668/12211 seeding of 24 values (random.cpp:508-520,
   prngcl_ranlux.cl:98-135).                                                */
static int rl_mlcg(int j)
{
    int k = j / 53668;
    j = 40014 * (j - k * 53668) - k * 12211;
    if (j < 0) j += RL_ICONS;
    return j;
}

void orc_ranlux_init_scalar(orc_ranlux *r, int jseed, int nskip)
{
    int j = jseed;
    for (int i = 0; i < 24; ++i) {
        j = rl_mlcg(j);
        r->s[i] = (float)(j % RL_TWO24) * RL_TWOM24;
    }
    r->carry = (r->s[23] == 0.0f) ? RL_TWOM24 : 0.0f;
    r->i24 = 23;
    r->j24 = 9;
    r->in24 = 0;
    r->nskip = nskip;
}

/* One subtract-with-borrow step at the ring pointers; returns the stored
   value (before the small-value fix).                                       */
static float rl_swb(orc_ranlux *r)
{
    float uni = r->s[r->j24] - r->s[r->i24] - r->carry;
    if (uni < 0.0f) {
        uni += 1.0f;
        r->carry = RL_TWOM24;
    } else {
        r->carry = 0.0f;
    }
    r->s[r->i24] = uni;
    if (r->i24 == 0) r->i24 = 24;
    if (r->j24 == 0) r->j24 = 24;
    r->i24--;
    r->j24--;
    return uni;
}

float orc_ranlux_next(orc_ranlux *r)
{
    if (r->in24 == 0) {
        r->in24 = 24;
        for (int k = 0; k < r->nskip; ++k) (void)rl_swb(r);
    }
    float uni = rl_swb(r);
    if (uni < RL_TWOM12) {
        uni = uni + RL_TWOM24 * r->s[r->j24];
        if (uni == 0.0f) uni = RL_TWOM24 * RL_TWOM24;
    }
    r->in24--;
    return uni;
}

void orc_ranlux_cpu_stream(unsigned randseries, int nskip, int count, float *out)
{
    orc_ranlux r;
    srand(randseries);
    orc_ranlux_init_scalar(&r, rand() % 2147483647, nskip);
    for (int i = 0; i < count; ++i) out[i] = orc_ranlux_next(&r);
}

/* Seeds: instance 0 = rand() % 2147483647 (random.cpp:508), the others
   rand() (random.cpp:464-465).  Table: seed_q (q=0..5) holds the logical
   values s[4q..4q+3]; float4 #6 = (23, 9, 6, carry) (prngcl_ranlux.cl:138-157),
   where .z = 6 marks "skip before the next quad".                          */
void orc_ranlux_seed_table(unsigned randseries, unsigned n, float *table)
{
    srand(randseries);
    for (unsigned g = 0; g < n; ++g) {
        int jseed = (g == 0) ? rand() % 2147483647 : rand();
        orc_ranlux r;
        orc_ranlux_init_scalar(&r, jseed, 24);
        for (int q = 0; q < 6; ++q)
            for (int c = 0; c < 4; ++c) table[4 * ((size_t)q * n + g) + c] = r.s[4 * q + c];
        float *car = table + 4 * ((size_t)6 * n + g);
        car[0] = 23.0f;
        car[1] = 9.0f;
        car[2] = 6.0f;
        car[3] = r.carry;
    }
}

/* The float4 kernel keeps the 24 values in six float4 registers and, after
   each skip, renames registers so that the ring restarts at (i24, j24) =
   (23, 9) (the RL_RANLUX2/3/4 and RL_skip_4/RL_skip_1 tails,
   prngcl_ranlux.cl:229-466).  Quad q of a block is generated from logical
   positions 23-4q .. 20-4q.  Restated here on the logical ring: load the
   table into a scalar ring, run the same SWB steps, rotate so the ring
   pointer is back at 23, store.                                            */
static void rl_table_load(const float *table, unsigned n, unsigned g, int nskip, orc_ranlux *r, int *quad)
{
    for (int q = 0; q < 6; ++q)
        for (int c = 0; c < 4; ++c) r->s[4 * q + c] = table[4 * ((size_t)q * n + g) + c];
    const float *car = table + 4 * ((size_t)6 * n + g);
    *quad = (int)car[2];
    r->carry = car[3];
    r->nskip = nskip;
    /* Ring pointers implied by the quad counter: quad k of the block starts
       at i24 = 23 - 4k (j24 = i24 - 14 mod 24).                             */
    int k = (*quad >= 6) ? 0 : *quad;
    r->i24 = 23 - 4 * k;
    r->j24 = (r->i24 + 10) % 24;
    r->in24 = (*quad >= 6) ? 0 : 24 - 4 * k;
}

static void rl_table_store(float *table, unsigned n, unsigned g, const orc_ranlux *r, int quad)
{
    for (int q = 0; q < 6; ++q)
        for (int c = 0; c < 4; ++c) table[4 * ((size_t)q * n + g) + c] = r->s[4 * q + c];
    float *car = table + 4 * ((size_t)6 * n + g);
    car[2] = (float)quad;
    car[3] = r->carry;
}

/* Skip nskip SWB steps, then rename the ring so i24 = 23 again.            */
static void rl_skip_and_rotate(orc_ranlux *r)
{
    for (int k = 0; k < r->nskip; ++k) (void)rl_swb(r);
    int shift = (23 - r->i24 + 24) % 24;     /* = nskip mod 24 */
    if (shift) {
        float tmp[24];
        for (int i = 0; i < 24; ++i) tmp[(i + shift) % 24] = r->s[i];
        memcpy(r->s, tmp, sizeof(tmp));
    }
    r->i24 = 23;
    r->j24 = 9;
}

void orc_ranlux_produce(float *table, unsigned n, unsigned samples, int nskip, float *quads)
{
    for (unsigned g = 0; g < n; ++g) {
        orc_ranlux r;
        int quad;
        rl_table_load(table, n, g, nskip, &r, &quad);
        for (unsigned i = 0; i < samples; ++i) {
            if (quad >= 6) {
                rl_skip_and_rotate(&r);
                quad = 0;
            }
            float *q = quads + 4 * ((size_t)i * n + g);
            for (int c = 0; c < 4; ++c) {
                float uni = rl_swb(&r);
                if (uni < RL_TWOM12) {
                    uni = uni + RL_TWOM24 * r.s[r.j24];
                    if (uni == 0.0f) uni = RL_TWOM24 * RL_TWOM24;
                }
                q[c] = uni;
            }
            quad++;
        }
        rl_table_store(table, n, g, &r, quad);
    }
}

/* ======================================================================== */
/* Lattice geometry -- sun_common.cl:41-66,351-380,455-513;                  */
/* suncpp/coord_work/coord_work.cpp:40-175                                   */
/* ======================================================================== */
unsigned orc_sites(const orc_lattice *L)
{
    return (unsigned)L->n[0] * L->n[1] * L->n[2] * L->n[3];
}

void orc_coords(const orc_lattice *L, unsigned site, int c[4])
{
    unsigned n2 = L->n[1], n3 = L->n[2], n4 = L->n[3];
    c[0] = (int)(site / (n2 * n3 * n4));
    unsigned r = site % (n2 * n3 * n4);
    c[3] = (int)(r / (n2 * n3));
    r %= n2 * n3;
    c[2] = (int)(r / n2);
    c[1] = (int)(r % n2);
}

unsigned orc_site(const orc_lattice *L, const int c[4])
{
    unsigned n2 = L->n[1], n3 = L->n[2], n4 = L->n[3];
    return (unsigned)c[1] + n2 * ((unsigned)c[2] + n3 * ((unsigned)c[3] + n4 * (unsigned)c[0]));
}

unsigned orc_neighbour(const orc_lattice *L, unsigned site, int dir, int sign)
{
    int c[4];
    orc_coords(L, site, c);
    c[dir] = (c[dir] + sign + L->n[dir]) % L->n[dir];
    return orc_site(L, c);
}

/* Thread i of a parity-p half-step updates site 2i or 2i+1, whichever has
   coordinate-sum parity p (sun_common.cl:351-380).                          */
unsigned orc_parity_site(const orc_lattice *L, unsigned i, int parity)
{
    int c[4];
    unsigned s = 2 * i;
    orc_coords(L, s, c);
    int p = (c[0] + c[1] + c[2] + c[3]) & 1;
    return (p == parity) ? s : s + 1;
}

/* ======================================================================== */
/* Small complex-matrix helpers (FP64).  Matrices are N x N complex,         */
/* row-major, [re, im] interleaved.  The product sums k in ascending order   */
/* with the term order of suncpp/su3/algebra_su3.cpp:118-140.                */
/* ======================================================================== */
static void mat_mul(int N, const double *a, const double *b, double *c)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < N; ++k) {
                double ar = a[2 * (i * N + k)], ai = a[2 * (i * N + k) + 1];
                double br = b[2 * (k * N + j)], bi = b[2 * (k * N + j) + 1];
                if (k == 0) {
                    re = -ai * bi + ar * br;
                    im = ar * bi + ai * br;
                } else {
                    re = re - ai * bi + ar * br;
                    im = im + ar * bi + ai * br;
                }
            }
            c[2 * (i * N + j)] = re;
            c[2 * (i * N + j) + 1] = im;
        }
}

static void mat_dagger(int N, const double *a, double *c)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            c[2 * (i * N + j)] = a[2 * (j * N + i)];
            c[2 * (i * N + j) + 1] = -a[2 * (j * N + i) + 1];
        }
}

static void mat_unit(int N, double *a)
{
    memset(a, 0, sizeof(double) * 2 * N * N);
    for (int i = 0; i < N; ++i) a[2 * (i * N + i)] = 1.0;
}

static double mat_retr(int N, const double *a)
{
    double t = a[0];
    for (int i = 1; i < N; ++i) t = t + a[2 * (i * N + i)];
    return t;
}

static double mat_imtr(int N, const double *a)
{
    double t = a[1];
    for (int i = 1; i < N; ++i) t = t + a[2 * (i * N + i) + 1];
    return t;
}

static inline double *LINK(const orc_lattice *L, double *links, unsigned site, int dir)
{
    return links + (size_t)2 * L->sun * L->sun * ((size_t)site * 4 + dir);
}
static inline const double *CLINK(const orc_lattice *L, const double *links, unsigned site, int dir)
{
    return links + (size_t)2 * L->sun * L->sun * ((size_t)site * 4 + dir);
}

/* Gram-Schmidt on rows 1-2 and row 3 = conj(row1 x row2)
   (suncpp/su3/algebra_su3.cpp:185-226).                                     */
void orc_gram_schmidt3(double *a)
{
    double *u = a, *v = a + 6, *w = a + 12;
    double nu = sqrt(u[0] * u[0] + u[2] * u[2] + u[4] * u[4] + u[1] * u[1] + u[3] * u[3] + u[5] * u[5]);
    double U[6];
    for (int k = 0; k < 6; ++k) U[k] = u[k] / nu;
    /* sp = sum_k v_k * conj(U_k), accumulated as (v1 U1*) + (v2 U2*) + (v3 U3*) */
    double spr = 0.0, spi = 0.0;
    for (int k = 0; k < 3; ++k) {
        double vr = v[2 * k], vi = v[2 * k + 1], ur = U[2 * k], ui = -U[2 * k + 1];
        double pr = vr * ur - vi * ui, pi = vr * ui + vi * ur;
        if (k == 0) {
            spr = pr;
            spi = pi;
        } else {
            spr = spr + pr;
            spi = spi + pi;
        }
    }
    double Vv[6];
    for (int k = 0; k < 3; ++k) {
        double ur = U[2 * k], ui = U[2 * k + 1];
        Vv[2 * k] = v[2 * k] - (ur * spr - ui * spi);
        Vv[2 * k + 1] = v[2 * k + 1] - (ur * spi + ui * spr);
    }
    double nv = sqrt(Vv[0] * Vv[0] + Vv[2] * Vv[2] + Vv[4] * Vv[4] + Vv[1] * Vv[1] + Vv[3] * Vv[3] + Vv[5] * Vv[5]);
    for (int k = 0; k < 6; ++k) Vv[k] /= nv;
    for (int k = 0; k < 6; ++k) {
        u[k] = U[k];
        v[k] = Vv[k];
    }
    /* matrix_reconstruct (algebra_su3.cpp:219-226): w = conj(u x v) */
    w[0] = u[5] * v[3] - u[4] * v[2] - u[3] * v[5] + u[2] * v[4];
    w[1] = u[4] * v[3] + u[5] * v[2] - u[2] * v[5] - u[3] * v[4];
    w[2] = -u[5] * v[1] + u[4] * v[0] + u[1] * v[5] - u[0] * v[4];
    w[3] = -u[4] * v[1] - u[5] * v[0] + u[0] * v[5] + u[1] * v[4];
    w[4] = u[3] * v[1] - u[2] * v[0] - u[1] * v[3] + u[0] * v[2];
    w[5] = u[2] * v[1] + u[3] * v[0] - u[0] * v[3] - u[1] * v[2];
}

/* SU(2) normalisation (su2_update_cl.cl:149-156): divide the quaternion
   (Re u1, Re u2, Im u1, Im u2) by its norm and rebuild row 2.               */
void orc_normalize2(double *a)
{
    double n = sqrt(a[0] * a[0] + a[2] * a[2] + a[1] * a[1] + a[3] * a[3]);
    a[0] /= n; a[1] /= n; a[2] /= n; a[3] /= n;
    a[4] = -a[2]; a[5] = a[3]; a[6] = a[0]; a[7] = -a[1];
}

void orc_reunitarize(const orc_lattice *L, double *links)
{
    unsigned V = orc_sites(L);
    for (unsigned s = 0; s < V; ++s)
        for (int d = 0; d < 4; ++d) {
            if (L->sun == 3) orc_gram_schmidt3(LINK(L, links, s, d));
            else orc_normalize2(LINK(L, links, s, d));
        }
}

/* ======================================================================== */
/* Starts -- suncpp/sunh.h:112-131; algebra_su3.cpp:64-80; algebra_su2.cpp:54-64 */
/* ======================================================================== */
void orc_init_cold(const orc_lattice *L, double *links)
{
    unsigned V = orc_sites(L);
    for (unsigned s = 0; s < V; ++s)
        for (int d = 0; d < 4; ++d) mat_unit(L->sun, LINK(L, links, s, d));
}

void orc_init_gid(const orc_lattice *L, double *links)
{
    unsigned V = orc_sites(L);
    double sites = (double)V;
    for (unsigned s = 0; s < V; ++s)
        for (int d = 0; d < 4; ++d) {
            double *a = LINK(L, links, s, d);
            double g = (double)s, f = (d + 1) / 1000.0;
            if (L->sun == 3) {
                a[0]  = sin((f + 10.0 / sites) * g) / 2.0;
                a[1]  = cos((f - 10.0 / sites) * g) / 3.0;
                a[2]  = sin((f - 18.0 / sites) * g) / 4.0;
                a[3]  = cos((f + 12.0 / sites) * g) / 4.0;
                a[4]  = sin((f - 12.0 / sites) * g) / 3.0;
                a[5]  = cos((f - 15.0 / sites) * g) / 2.0;
                a[6]  = sin((f + 15.0 / sites) * g) / 3.0;
                a[7]  = cos((f - 11.0 / sites) * g) / 4.0;
                a[8]  = sin((f - 11.0 / sites) * g) / 3.0;
                a[9]  = cos((f + 16.0 / sites) * g) / 2.0;
                a[10] = sin((f - 16.0 / sites) * g) / 3.0;
                a[11] = cos((f - 18.0 / sites) * g) / 2.0;
                orc_gram_schmidt3(a);
            } else {
                a[0] = sin((f + 10.0 / sites) * g) / 2.0;   /* Re u1 */
                a[3] = cos((f - 10.0 / sites) * g) / 3.0;   /* Im u2 */
                a[2] = sin((f - 18.0 / sites) * g) / 4.0;   /* Re u2 */
                a[1] = sqrt(1 - a[0] * a[0] - a[3] * a[3] - a[2] * a[2]);
                a[4] = -a[2];
                a[5] = a[3];
                a[6] = a[0];
                a[7] = -a[1];
            }
        }
}

/* Haar-random start from per-stream uniforms (new work: the reference's
   lattice_init_hot_* is not Haar, su2_update_cl.cl:110-146,
   su3_update_cl.cl:165-225).  Stream i (= compact index) fills the links of
   its two sites in the order parity 0, 1 and direction 0..3; each link takes
   Q quads (Q = 1 for SU(2), 3 for SU(3)): quad k of stream i is
   quads[(k * stride + i) * 4 + c].  Gaussians by Box-Muller in double,
   g = sqrt(-2 ln t) (cos 2 pi u2, sin 2 pi u2), t = 1 - u1 (t = 2^-33 if
   u1 = 1), from the pairs (x, y) and (z, w) of each quad.
   SU(2): (Re u11, Im u11, Re u12, Im u12) = g / |g|, uniform on S^3.
   SU(3): rows 1, 2 = (g0 + i g1, g2 + i g3, g4 + i g5), (g6 + i g7, ...),
   Gram-Schmidt, row 3 = conj(row 1 x row 2): a uniform orthonormal 2-frame
   completed to SU(3), which is Haar (SU(3) acts simply transitively on
   orthonormal 2-frames of C^3).                                            */
static void orc_box_muller(float u1, float u2, double *g0, double *g1)
{
    double t = 1.0 - (double)u1;
    if (t <= 0.0) t = 1.1641532182693481e-10;   /* 2^-33 */
    double r = sqrt(-2.0 * log(t));
    double th = 6.2831853071795864769 * (double)u2;
    *g0 = r * cos(th);
    *g1 = r * sin(th);
}

void orc_init_hot(const orc_lattice *L, double *links, const float *quads, unsigned stride)
{
    unsigned vh = orc_sites(L) / 2;
    int Q = (L->sun == 3) ? 3 : 1;
    for (unsigned i = 0; i < vh; ++i)
        for (int p = 0; p < 2; ++p) {
            unsigned s = orc_parity_site(L, i, p);
            for (int d = 0; d < 4; ++d) {
                double g[12];
                for (int q = 0; q < Q; ++q) {
                    const float *u = quads + 4 * ((size_t)((p * 4 + d) * Q + q) * stride + i);
                    orc_box_muller(u[0], u[1], &g[4 * q], &g[4 * q + 1]);
                    orc_box_muller(u[2], u[3], &g[4 * q + 2], &g[4 * q + 3]);
                }
                double *a = LINK(L, links, s, d);
                if (L->sun == 3) {
                    for (int k = 0; k < 12; ++k) a[k] = g[k];
                    orc_gram_schmidt3(a);
                } else {
                    double n = 1.0 / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
                    a[0] = g[0] * n;
                    a[1] = g[1] * n;
                    a[2] = g[2] * n;
                    a[3] = g[3] * n;
                    a[4] = -a[2];
                    a[5] = a[3];
                    a[6] = a[0];
                    a[7] = -a[1];
                }
            }
        }
}

/* ======================================================================== */
/* Staple -- suncpp/Update/sun_update.h:44-77 (term and summation order)     */
/* ======================================================================== */
void orc_staple(const orc_lattice *L, const double *links, unsigned site, int mu, double *A)
{
    int N = L->sun;
    double m1[18], m2[18], m3[18], t[18], p[18];
    memset(A, 0, sizeof(double) * 2 * N * N);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        unsigned xpm = orc_neighbour(L, site, mu, +1);
        unsigned xpn = orc_neighbour(L, site, nu, +1);
        memcpy(m1, CLINK(L, links, xpm, nu), sizeof(double) * 2 * N * N);
        mat_dagger(N, CLINK(L, links, xpn, mu), m2);
        mat_dagger(N, CLINK(L, links, site, nu), m3);
        mat_mul(N, m1, m2, t);
        mat_mul(N, t, m3, p);
        for (int k = 0; k < 2 * N * N; ++k) A[k] = A[k] + p[k];

        unsigned xmn = orc_neighbour(L, site, nu, -1);
        unsigned xmnpm = orc_neighbour(L, xmn, mu, +1);
        memcpy(m3, CLINK(L, links, xmn, nu), sizeof(double) * 2 * N * N);
        mat_dagger(N, CLINK(L, links, xmn, mu), m2);
        mat_dagger(N, CLINK(L, links, xmnpm, nu), m1);
        mat_mul(N, m1, m2, t);
        mat_mul(N, t, m3, p);
        for (int k = 0; k < 2 * N * N; ++k) A[k] = A[k] + p[k];
    }
}

double orc_link_action(const orc_lattice *L, const double *links, unsigned site, int mu)
{
    double A[18], W[18];
    orc_staple(L, links, site, mu, A);
    mat_mul(L->sun, CLINK(L, links, site, mu), A, W);
    return mat_retr(L->sun, W);
}

/* ======================================================================== */
/* Kennedy-Pendleton SU(2) heatbath -- suncpp/su2/update_su2.cpp:40-126 with */
/* the GPU quad-consumption rule of su3_update_cl.cl:568-669 for QUADS.      */
/*   b: 2x2 complex block (8 doubles) in; out: new 2x2 (8 doubles).          */
/*   Returns 1 if accepted, 0 if all NHIT tries were rejected.               */
/* ======================================================================== */
typedef struct {
    orc_rng *rng;
    unsigned thread;    /* QUADS: thread index i                              */
    unsigned k;         /* QUADS: next quad of this thread                    */
    unsigned gid;       /* GID mode: site index                               */
    unsigned sites;     /* GID mode: lattice volume                           */
} kp_ctx;

static void kp_fetch(kp_ctx *c, float r[4], int n)
{
    orc_rng *g = c->rng;
    if (g->kind == ORC_RNG_SEQ) {
        for (int i = 0; i < n; ++i)
            r[i] = (g->seq_generator == 0) ? orc_ranlux_next(g->ranlux) : orc_pm_one_cpu(g->pm_seed);
    } else {
        const float *q = g->quads + 4 * ((size_t)c->thread + (size_t)c->k * g->stride);
        for (int i = 0; i < 4; ++i) r[i] = q[i];
        c->k++;
    }
}

static int kp_heatbath(const double *b, double beta, int nhit, int mode, kp_ctx *c, double *out)
{
    double Mx = b[0] + b[6];
    double My = b[1] - b[7];
    double Mz = b[2] - b[4];
    double Mw = b[3] + b[5];
    double det = sqrt(Mx * Mx + My * My + Mz * Mz + Mw * Mw);
    double bdet = beta * det;
    double aH[8];
    aH[0] = Mx / det;
    aH[1] = -My / det;
    aH[2] = -Mz / det;
    aH[3] = -Mw / det;
    aH[6] = aH[0];
    aH[7] = -aH[1];
    aH[4] = -aH[2];
    aH[5] = aH[3];

    double cm[8];
    if (mode == ORC_OR) {
        /* Overrelaxation: the reflection R = aH * aH keeps ReTr(R W) fixed. */
        memcpy(cm, aH, sizeof(cm));
        mat_mul(2, cm, aH, out);
        return 1;
    }

    int gidm = (c->rng->kind == ORC_RNG_GID);
    int flag = 0, i = 0;
    double delta = 0.0;
    float rnd[4];
    double gid = (double)c->gid, fs = (double)c->sites;
    while (i < nhit && !flag) {
        if (gidm) {
            rnd[0] = fabsf((float)sin((0.005 * (1.0 + nhit) + 270.0 / fs) * gid));
            rnd[1] = fabsf((float)cos((0.005 * (1.0 + nhit) + 60.0 / fs) * gid));
            rnd[2] = fabsf((float)sin((0.005 * (1.0 + nhit) - 150.0 / fs) * gid));
            rnd[3] = fabsf((float)cos((0.005 * (1.0 + nhit) - 380.0 / fs) * gid));
        } else {
            kp_fetch(c, rnd, 4);
        }
        double cosrnd = cos(ORC_PI2 * rnd[1]);
        if (!gidm) {
            delta = -(log(1.0 - rnd[0]) + cosrnd * cosrnd * log(1.0 - rnd[2])) / bdet;
            float w2 = rnd[3] * rnd[3];
            if ((double)w2 <= (1.0 - 0.5 * delta)) flag = 1;
        } else {
            delta = cosrnd * cosrnd / bdet;
            flag = 1;
        }
        i++;
    }
    if (!flag) {
        if (c->rng->kind == ORC_RNG_QUADS) c->k++;   /* su3_update_cl.cl:665-668 */
        return 0;
    }
    if (gidm) {
        rnd[0] = fabsf((float)cos((0.08 - 270.0 / fs) * gid));
        rnd[1] = fabsf((float)sin((0.08 - 60.0 / fs) * gid));
    } else {
        kp_fetch(c, rnd, 2);
    }
    double cosal = 1.0 - delta;
    double costh = 2.0 * rnd[0] - 1.0;
    double sinth = sqrt(1.0 - costh * costh);
    double sinal = sqrt(1.0 - cosal * cosal);
    double phi = ORC_PI2 * rnd[1];
    double sinphi = sin(phi), cosphi = cos(phi);
    cm[0] = cosal;
    cm[1] = sinal * costh;
    cm[2] = sinal * sinth * sinphi;
    cm[3] = sinal * sinth * cosphi;
    cm[6] = cm[0];
    cm[7] = -cm[1];
    cm[4] = -cm[2];
    cm[5] = cm[3];
    mat_mul(2, cm, aH, out);
    return 1;
}

/* ======================================================================== */
/* Link updates                                                              */
/* ======================================================================== */
/* Subgroup index pairs of the Cabibbo-Marinari loop, in the reference order
   (suncpp/su3/update_su3.cpp:46-109; su3_update_cl.cl:687-771).            */
static const int CM_PAIRS[3][2] = {{0, 1}, {0, 2}, {1, 2}};

/* One SU(3) link: U <- (R3 R2 R1) U with R_k from the heatbath (or the
   overrelaxation reflection) in subgroup k, then Gram-Schmidt. */
static int update_su3(double *U, const double *A, double beta, int nhit, int mode, kp_ctx *c)
{
    double U0[18], x0[18], Vg[18], Un[18], r[8], rn[8];
    int rejected = 0;
    memcpy(U0, U, sizeof(U0));
    for (int sg = 0; sg < 3; ++sg) {
        int a = CM_PAIRS[sg][0], b = CM_PAIRS[sg][1];
        mat_mul(3, U0, A, x0);
        r[0] = x0[2 * (a * 3 + a)];  r[1] = x0[2 * (a * 3 + a) + 1];
        r[2] = x0[2 * (a * 3 + b)];  r[3] = x0[2 * (a * 3 + b) + 1];
        r[4] = x0[2 * (b * 3 + a)];  r[5] = x0[2 * (b * 3 + a) + 1];
        r[6] = x0[2 * (b * 3 + b)];  r[7] = x0[2 * (b * 3 + b) + 1];
        if (!kp_heatbath(r, beta, nhit, mode, c, rn)) {
            rejected++;
            memset(rn, 0, sizeof(rn));      /* lattice_unity (update_su2.cpp:123-124) */
            rn[0] = 1.0;
            rn[6] = 1.0;
        }
        mat_unit(3, Vg);
        Vg[2 * (a * 3 + a)] = rn[0];  Vg[2 * (a * 3 + a) + 1] = rn[1];
        Vg[2 * (a * 3 + b)] = rn[2];  Vg[2 * (a * 3 + b) + 1] = rn[3];
        Vg[2 * (b * 3 + a)] = rn[4];  Vg[2 * (b * 3 + a) + 1] = rn[5];
        Vg[2 * (b * 3 + b)] = rn[6];  Vg[2 * (b * 3 + b) + 1] = rn[7];
        mat_mul(3, Vg, U0, Un);
        memcpy(U0, Un, sizeof(U0));
    }
    orc_gram_schmidt3(U0);
    memcpy(U, U0, sizeof(U0));
    return rejected;
}

/* One SU(2) link (suncpp/su2/update_su2.cpp with final = 1: a rejected
   update keeps the old link).  Overrelaxation reflects U through the staple
   direction: U <- (W^ dagger)^2 U with W = U A.                             */
static int update_su2(double *U, const double *A, double beta, int nhit, int mode, kp_ctx *c)
{
    double rn[8];
    if (mode == ORC_OR) {
        double W[8], R[8], Un[8];
        mat_mul(2, U, A, W);
        kp_heatbath(W, beta, nhit, ORC_OR, c, R);
        mat_mul(2, R, U, Un);
        memcpy(U, Un, sizeof(Un));
        return 0;
    }
    if (!kp_heatbath(A, beta, nhit, mode, c, rn)) return 1;
    memcpy(U, rn, sizeof(rn));
    return 0;
}

int orc_halfstep(const orc_lattice *L, double *links, int parity, int mu, int mode, orc_rng *rng)
{
    unsigned V = orc_sites(L);
    double beta = L->beta / L->sun;            /* sun_update.h:95; suncl.cpp:3940 */
    int N = L->sun, rejected = 0;
    double A[18];
    for (unsigned i = 0; i < V / 2; ++i) {
        unsigned s = orc_parity_site(L, i, parity);
        orc_staple(L, links, s, mu, A);
        kp_ctx c = {rng, i, 0, s, V};
        double *U = LINK(L, links, s, mu);
        if (N == 3) rejected += update_su3(U, A, beta, L->nhit, mode, &c);
        else rejected += update_su2(U, A, beta, L->nhit, mode, &c);
    }
    return rejected;
}

/* The same half-step with the staples read from `src` (a copy of the
   links taken before it) and the updated links written into `links`: the
   double-buffered form the GPU's shadow mode runs.  For parity = 2 the
   half-step takes the sites of both parities in one pass (a planted read
   of links the same pass writes, for the tests of the race check).        */
int orc_halfstep_src(const orc_lattice *L, double *links, const double *src, int parity, int mu, int mode,
                     orc_rng *rng)
{
    unsigned V = orc_sites(L);
    double beta = L->beta / L->sun;
    int N = L->sun, rejected = 0;
    double A[18];
    const int np = parity == 2 ? 2 : 1;
    for (int k = 0; k < np; ++k) {
        const int par = parity == 2 ? 1 - k : parity;
        for (unsigned i = 0; i < V / 2; ++i) {
            unsigned s = orc_parity_site(L, i, par);
            orc_staple(L, src ? src : links, s, mu, A);
            kp_ctx c = {rng, i, 0, s, V};
            double *U = LINK(L, links, s, mu);
            if (N == 3) rejected += update_su3(U, A, beta, L->nhit, mode, &c);
            else rejected += update_su2(U, A, beta, L->nhit, mode, &c);
        }
    }
    return rejected;
}

int orc_sweep(const orc_lattice *L, double *links, int mode, orc_rng *rng)
{
    int rejected = 0;
    for (int parity = 1; parity >= 0; --parity)        /* odd first (suncpp.h:159-167) */
        for (int mu = 0; mu < 4; ++mu) rejected += orc_halfstep(L, links, parity, mu, mode, rng);
    return rejected;
}

/* ======================================================================== */
/* Measurements                                                              */
/* ======================================================================== */
/* Plaquette sums, planes as in suncpp/Measurements/Plq.h:47-94 and
   sun_measurements_cl.cl:1178-1309: spatial (x,y),(x,z),(y,z); temporal
   (x,t),(y,t),(z,t).  ReTr[U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+].           */
static double plaq(const orc_lattice *L, const double *links, unsigned s, int a, int b)
{
    int N = L->sun;
    double m3[18], m4[18], t1[18], t2[18], t3[18];
    unsigned xa = orc_neighbour(L, s, a, +1), xb = orc_neighbour(L, s, b, +1);
    mat_dagger(N, CLINK(L, links, xb, a), m3);
    mat_dagger(N, CLINK(L, links, s, b), m4);
    mat_mul(N, CLINK(L, links, s, a), CLINK(L, links, xa, b), t1);
    mat_mul(N, t1, m3, t2);
    mat_mul(N, t2, m4, t3);
    return mat_retr(N, t3);
}

void orc_plaquette(const orc_lattice *L, const double *links, double plq[2])
{
    unsigned V = orc_sites(L);
    double sp = 0.0, tm = 0.0;
    for (unsigned s = 0; s < V; ++s) {
        sp += plaq(L, links, s, 0, 1);
        sp += plaq(L, links, s, 0, 2);
        sp += plaq(L, links, s, 1, 2);
        tm += plaq(L, links, s, 0, 3);
        tm += plaq(L, links, s, 1, 3);
        tm += plaq(L, links, s, 2, 3);
    }
    plq[0] = sp;
    plq[1] = tm;
}

/* Per-site spatial and temporal plaquette sums (the per-site values the
   action per slice adds up, lattice_action_diff_x/y/z,
   sun_measurements_cl.cl:186-615): out[2*site + 0/1], site order of
   orc_site.                                                                 */
void orc_plaquette_sites(const orc_lattice *L, const double *links, double *out)
{
    unsigned V = orc_sites(L);
    for (unsigned s = 0; s < V; ++s) {
        out[2 * s] = plaq(L, links, s, 0, 1) + plaq(L, links, s, 0, 2) + plaq(L, links, s, 1, 2);
        out[2 * s + 1] = plaq(L, links, s, 0, 3) + plaq(L, links, s, 1, 3) + plaq(L, links, s, 2, 3);
    }
}

/* F_munu (lattice_measurement_plq with FMUNU / F0MU,
   sun_measurements_cl.cl:710-1173): for each site and plane the plaquette
   P = U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+ and F = -i Tr(P G) / 2 for the
   generator G, written out element by element as trace_lambda1..8
   (su3_measurements_cl.cl:357-460) and trace_sigma1..3
   (su2_measurements_cl.cl:89-123; SU(2) imaginary part 0 and no 1/2: its
   P is stored as the first row (u1, u2) only).  Planes (x,y), (x,z), (y,z)
   or, efield = 1, (x,t), (y,t), (z,t).
   out[6*g + 2*plane + (0 re, 1 im)] summed over sites, g = 0 for gen1.    */
static void plaq_matrix(const orc_lattice *L, const double *links, unsigned s, int a, int b, double *P)
{
    int N = L->sun;
    double m3[18], m4[18], t1[18], t2[18];
    unsigned xa = orc_neighbour(L, s, a, +1), xb = orc_neighbour(L, s, b, +1);
    mat_dagger(N, CLINK(L, links, xb, a), m3);
    mat_dagger(N, CLINK(L, links, s, b), m4);
    mat_mul(N, CLINK(L, links, s, a), CLINK(L, links, xa, b), t1);
    mat_mul(N, t1, m3, t2);
    mat_mul(N, t2, m4, P);
}

static void trace_gen3(const double *P, int k, double *re, double *im)
{
#define RE(r, c) P[2 * ((r) * 3 + (c))]
#define IM(r, c) P[2 * ((r) * 3 + (c)) + 1]
    double a = 0.0, b = 0.0;
    switch (k) {
        case 1: a = IM(0, 1) + IM(1, 0); b = -RE(0, 1) - RE(1, 0); break;
        case 2: a = RE(0, 1) - RE(1, 0); b = IM(0, 1) - IM(1, 0); break;
        case 3: a = IM(0, 0) - IM(1, 1); b = -RE(0, 0) + RE(1, 1); break;
        case 4: a = IM(0, 2) + IM(2, 0); b = -RE(0, 2) - RE(2, 0); break;
        case 5: a = RE(0, 2) - RE(2, 0); b = IM(0, 2) - IM(2, 0); break;
        case 6: a = IM(1, 2) + IM(2, 1); b = -RE(1, 2) - RE(2, 1); break;
        case 7: a = RE(1, 2) - RE(2, 1); b = IM(1, 2) - IM(2, 1); break;
        default:
            a = (IM(0, 0) + IM(1, 1) - 2.0 * IM(2, 2)) * 0.57735026918962576451;
            b = (-RE(0, 0) - RE(1, 1) + 2.0 * RE(2, 2)) * 0.57735026918962576451;
    }
#undef RE
#undef IM
    *re = a / 2;
    *im = b / 2;
}

static void trace_gen2(const double *P, int k, double *re, double *im)
{
    /* first row u1 = P[0][0], u2 = P[0][1] */
    *re = (k == 1) ? P[3] : (k == 2) ? P[2] : P[1];
    *im = 0.0;
}

void orc_fmunu(const orc_lattice *L, const double *links, int efield, int gen1, int gen2, double out[12])
{
    static const int pa_h[3] = {0, 0, 1}, pb_h[3] = {1, 2, 2};
    static const int pa_e[3] = {0, 1, 2}, pb_e[3] = {3, 3, 3};
    const int *pa = efield ? pa_e : pa_h, *pb = efield ? pb_e : pb_h;
    const int gens[2] = {gen1, gen2};
    unsigned V = orc_sites(L);
    double P[18];
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    for (unsigned s = 0; s < V; ++s)
        for (int pl = 0; pl < 3; ++pl) {
            plaq_matrix(L, links, s, pa[pl], pb[pl], P);
            for (int g = 0; g < 2; ++g) {
                double re, im;
                if (L->sun == 3) trace_gen3(P, gens[g], &re, &im);
                else trace_gen2(P, gens[g], &re, &im);
                out[6 * g + 2 * pl] += re;
                out[6 * g + 2 * pl + 1] += im;
            }
        }
}

/* R x T Wilson loop summed over sites and the three spatial directions d
   (su3cpu.cpp:1435-1522, wilson_loop.cl:55-465): W = B R T^+ L^+ with
   B = U_t(x) ... U_t(x+(T-1)t), R = U_d(x+Tt) ... U_d(x+Tt+(R-1)d),
   T = U_t(x+Rd) ... U_t(x+Rd+(T-1)t), L = U_d(x) ... U_d(x+(R-1)d);
   returns sum ReTr W (the reference divides by 3V).                        */
static void line_product(const orc_lattice *L, const double *links, unsigned site, int dir, int len, double *P)
{
    int N = L->sun;
    double T[18];
    memcpy(P, CLINK(L, links, site, dir), sizeof(double) * 2 * N * N);
    for (int k = 1; k < len; ++k) {
        site = orc_neighbour(L, site, dir, +1);
        mat_mul(N, P, CLINK(L, links, site, dir), T);
        memcpy(P, T, sizeof(double) * 2 * N * N);
    }
}

double orc_wilson(const orc_lattice *L, const double *links, int R, int T)
{
    int N = L->sun;
    unsigned V = orc_sites(L);
    double sum = 0.0;
    double B[18], Rm[18], Tm[18], Lm[18], m1[18], m2[18], m3[18];
    for (unsigned x = 0; x < V; ++x) {
        line_product(L, links, x, 3, T, B);
        unsigned xt = x;
        for (int k = 0; k < T; ++k) xt = orc_neighbour(L, xt, 3, +1);
        for (int d = 0; d < 3; ++d) {
            unsigned xr = x;
            for (int k = 0; k < R; ++k) xr = orc_neighbour(L, xr, d, +1);
            line_product(L, links, xt, d, R, Rm);
            line_product(L, links, xr, 3, T, Tm);
            line_product(L, links, x, d, R, Lm);
            mat_mul(N, B, Rm, m1);
            mat_dagger(N, Tm, m2);
            mat_mul(N, m1, m2, m3);
            mat_dagger(N, Lm, m2);
            mat_mul(N, m3, m2, m1);
            sum += mat_retr(N, m1);
        }
    }
    return sum;
}

/* Polyakov loop: P(x) = U_T(x,0) U_T(x,1) ... U_T(x,N4-1), traces in FP64
   (polyakov.cl:54-141; su3cpu.cpp:1383-1433).                               */
void orc_polyakov(const orc_lattice *L, const double *links, double poly[4])
{
    int N = L->sun;
    double P[18], T[18];
    double re = 0.0, im = 0.0, p2 = 0.0, p4 = 0.0;
    for (int x = 0; x < L->n[0]; ++x)
        for (int z = 0; z < L->n[2]; ++z)
            for (int y = 0; y < L->n[1]; ++y) {
                int c[4] = {x, y, z, 0};
                unsigned s = orc_site(L, c);
                memcpy(P, CLINK(L, links, s, 3), sizeof(double) * 2 * N * N);
                for (int t = 1; t < L->n[3]; ++t) {
                    c[3] = t;
                    mat_mul(N, P, CLINK(L, links, orc_site(L, c), 3), T);
                    memcpy(P, T, sizeof(double) * 2 * N * N);
                }
                double r = mat_retr(N, P), i = mat_imtr(N, P);
                double a2 = r * r + i * i;
                re += r;
                im += i;
                p2 += a2;
                p4 += a2 * a2;
            }
    poly[0] = re;
    poly[1] = im;
    poly[2] = p2;
    poly[3] = p4;
}
