/*
 * oracle.h -- CPU restatement of QCDGPU's SU(N) Monte Carlo update and
 * measurement path, used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check or to time the
 * reference algorithm on the host.  The product (libqcdhip.so) never links,
 * loads or calls it.
 *
 * Everything is FP64 on the host except the pseudo-random numbers, which are
 * float32 exactly as in the reference.  Link storage is the canonical host
 * layout shared with the product's import/export calls:
 *     links[((site * 4 + dir) * N + row) * N + col][re, im]     (doubles)
 * with site = y + z*N2 + t*N2*N3 + x*N2*N3*N4 (reference sun_common.cl:62-66)
 * and dims n[0..3] = (N1, N2, N3, N4) = extents along (x, y, z, t).
 *
 * Reference anchors for each function are in oracle.c.
 */
#ifndef QCD_ORACLE_H
#define QCD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- glibc */
void     orc_srand(unsigned seed);
int      orc_rand(void);
void     orc_rand_fill(unsigned seed, int count, int *out);

/* ---------------------------------------------------------------- XOR128 */
/* One Marsaglia xorshift128 step on (x,y,z,w); returns the new w. */
uint32_t orc_xor128_step(uint32_t s[4]);
/* Reference seed table for `n_instances` instances (random.cpp:235,647-659). */
void     orc_xor128_seed_table(unsigned randseries, unsigned n_instances, uint32_t *table);
/* Reference "xor128" kernel: fills quads[(i * n + g) * 4 + c], advances table. */
void     orc_xor128_produce(uint32_t *table, unsigned n, unsigned samples, float *quads);

/* ---------------------------------------------------------------- Park-Miller */
uint32_t orc_pm_step(uint32_t s);
void     orc_pm_seed_table(unsigned randseries, unsigned n_instances, uint32_t *table);
void     orc_pm_produce(uint32_t *table, unsigned n, unsigned samples, float *quads);
/* CPU twin (random.cpp:896-910): returns the float, advances *seed. */
float    orc_pm_one_cpu(int32_t *seed);

/* ---------------------------------------------------------------- RANLUX */
typedef struct {
    float   s[24];     /* lagged values, logical index as in the reference  */
    int     i24, j24;  /* scalar ring pointers                               */
    int     in24;      /* outputs left in the current 24-block               */
    float   carry;
    int     nskip;     /* numbers discarded per 24-block (24/48/97/223/389)  */
} orc_ranlux;

int      orc_ranlux_nskip(int luxury);            /* 0..4 -> 24,48,97,223,389 */
void     orc_ranlux_init_scalar(orc_ranlux *r, int jseed, int nskip);
float    orc_ranlux_next(orc_ranlux *r);
/* The reference's CPU generator after srand(randseries) (random.cpp:504-520). */
void     orc_ranlux_cpu_stream(unsigned randseries, int nskip, int count, float *out);
/* Reference seed table: 7 float4 per instance, table[(q * n + g) * 4 + c]. */
void     orc_ranlux_seed_table(unsigned randseries, unsigned n_instances, float *table);
/* Reference "ranlux" kernel over the table (prngcl_ranlux.cl:534-563). */
void     orc_ranlux_produce(float *table, unsigned n, unsigned samples, int nskip, float *quads);

/* ---------------------------------------------------------------- lattice */
typedef struct {
    int      sun;        /* 2 or 3                                           */
    int      n[4];       /* N1..N4                                           */
    double   beta;       /* BETA; the update uses BETA / sun                 */
    int      nhit;       /* NHIT                                             */
} orc_lattice;

unsigned orc_sites(const orc_lattice *L);
void     orc_coords(const orc_lattice *L, unsigned site, int c[4]);
unsigned orc_site(const orc_lattice *L, const int c[4]);
unsigned orc_neighbour(const orc_lattice *L, unsigned site, int dir, int sign);
unsigned orc_parity_site(const orc_lattice *L, unsigned i, int parity);

void     orc_init_cold(const orc_lattice *L, double *links);
void     orc_init_gid(const orc_lattice *L, double *links);
/* Haar start from per-stream quads: quads[(k * stride + i) * 4 + c] is quad
   k of stream i (compact index); k runs over parity, direction and the
   Q = 1 (SU(2)) or 3 (SU(3)) quads of a link. */
void     orc_init_hot(const orc_lattice *L, double *links, const float *quads, unsigned stride);

/* Random-number sources for the update. */
enum { ORC_RNG_GID = 0, ORC_RNG_SEQ = 1, ORC_RNG_QUADS = 2 };
typedef struct {
    int          kind;
    /* ORC_RNG_SEQ: sequential scalar generator, 4 numbers per KP try and 2
       after acceptance (suncpp/su2/update_su2.cpp:77,101)                   */
    int          seq_generator;   /* 0 = RANLUX (ranlux), 1 = PM (pm_seed)   */
    orc_ranlux  *ranlux;
    int32_t     *pm_seed;
    /* ORC_RNG_QUADS: the reference GPU buffer semantics: thread i reads quad
       quads[i + k * stride] (su3_update_cl.cl:613-617,644-648)             */
    const float *quads;
    unsigned long stride;
} orc_rng;

enum { ORC_HB = 0, ORC_OR = 1 };

void     orc_staple(const orc_lattice *L, const double *links, unsigned site, int mu, double *A);
/* One (parity, mu) half-step.  For ORC_RNG_QUADS the caller provides the
   quads of this half-step. Returns the number of rejected-all subgroups. */
int      orc_halfstep(const orc_lattice *L, double *links, int parity, int mu, int mode, orc_rng *rng);
int      orc_halfstep_src(const orc_lattice *L, double *links, const double *src, int parity, int mu, int mode,
                          orc_rng *rng);
/* One sweep in the reference order odd X..T, even X..T. For QUADS the
   caller-provided callback is not available, so only GID/SEQ are allowed. */
int      orc_sweep(const orc_lattice *L, double *links, int mode, orc_rng *rng);

void     orc_gram_schmidt3(double *u);            /* 18 doubles in place     */
void     orc_normalize2(double *u);               /* 8 doubles in place      */
void     orc_reunitarize(const orc_lattice *L, double *links);

/* Measurements (FP64): sums over sites, not normalised.
   plq[0] = sum of ReTr over the 3 spatial planes, plq[1] over the 3 temporal. */
void     orc_plaquette(const orc_lattice *L, const double *links, double plq[2]);
/* Per-site (spatial, temporal) plaquette sums, out[2*site + c].            */
void     orc_plaquette_sites(const orc_lattice *L, const double *links, double *out);
/* poly[0..3] = sum over spatial sites of ReTr P, ImTr P, |Tr P|^2, |Tr P|^4. */
/* Sum over sites and spatial directions of ReTr of the R x T Wilson loop
   (spatial extent R along x, y, z; temporal extent T along t).             */
double   orc_wilson(const orc_lattice *L, const double *links, int R, int T);
void     orc_polyakov(const orc_lattice *L, const double *links, double poly[4]);
/* F_munu sums: out[6*g + 2*plane + (re, im)], generators gen1 (g = 0) and
   gen2 (g = 1), H-field planes (x,y), (x,z), (y,z) or E-field (x,t), (y,t),
   (z,t) for efield = 1.                                                    */
void     orc_fmunu(const orc_lattice *L, const double *links, int efield, int gen1, int gen2, double out[12]);
/* Sum over links of ReTr(U A) (A = staple), the local action density. */
double   orc_link_action(const orc_lattice *L, const double *links, unsigned site, int mu);

#ifdef __cplusplus
}
#endif
#endif
