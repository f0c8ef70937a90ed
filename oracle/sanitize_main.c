/* Test infrastructure (the oracle's own memory and undefined-behaviour
   check): a driver that runs every family of oracle routines once --
   generators and their seed tables, the starts, heatbath and
   overrelaxation half-steps from each random-number source, the
   double-buffered half-step, reunitarisation and every measurement -- on
   small lattices whose extents all differ, so that an index that mixes two
   axes or runs past a buffer is hit.  Built with AddressSanitizer and
   UndefinedBehaviorSanitizer by tests/test_sanitize.py; it prints one
   checksum line per lattice and exits 0.  Not part of the product.        */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double checksum(const double *a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * (double)((i % 7) + 1);
    return s;
}

static int run_lattice(int sun, const int n[4], double beta)
{
    orc_lattice L = {sun, {n[0], n[1], n[2], n[3]}, beta, 10};
    const unsigned V = orc_sites(&L);
    const size_t nl = (size_t)V * 4 * sun * sun * 2;
    double *links = malloc(sizeof(double) * nl), *copy = malloc(sizeof(double) * nl);
    if (!links || !copy) return 1;

    /* starts */
    orc_init_gid(&L, links);
    orc_init_cold(&L, links);
    const unsigned vh = V / 2, stride = (vh + 63) / 64 * 64;
    const unsigned per_stream = 8 * (sun == 3 ? 3 : 1);   /* two parities x four directions x Q */
    float *quads = malloc(sizeof(float) * 4 * (size_t)per_stream * stride);
    uint32_t *xt = malloc(sizeof(uint32_t) * 4 * stride);
    if (!quads || !xt) return 1;
    orc_xor128_seed_table(54, stride, xt);
    orc_xor128_produce(xt, stride, per_stream, quads);
    orc_init_hot(&L, links, quads, stride);

    /* heatbath from the reference's quad buffer, sequential RANLUX, PM and
       GID numbers; overrelaxation; double-buffered half-steps */
    const unsigned s_quads = 1 + ((3 * stride * 11 + 63) / 64 * 64 - 1) / stride;
    float *hb = malloc(sizeof(float) * 4 * (size_t)s_quads * stride);
    if (!hb) return 1;
    orc_rng rq = {0};
    rq.kind = ORC_RNG_QUADS;
    rq.quads = hb;
    rq.stride = stride;
    for (int parity = 1; parity >= 0; --parity)
        for (int mu = 0; mu < 4; ++mu) {
            orc_xor128_produce(xt, stride, s_quads, hb);
            orc_halfstep(&L, links, parity, mu, ORC_HB, &rq);
        }
    orc_ranlux rl;
    orc_srand(7);
    orc_ranlux_init_scalar(&rl, orc_rand() % 2147483647, orc_ranlux_nskip(3));
    orc_rng rs = {0};
    rs.kind = ORC_RNG_SEQ;
    rs.seq_generator = 0;
    rs.ranlux = &rl;
    orc_sweep(&L, links, ORC_HB, &rs);
    int32_t pm = 12345;
    orc_rng rp = {0};
    rp.kind = ORC_RNG_SEQ;
    rp.seq_generator = 1;
    rp.pm_seed = &pm;
    orc_sweep(&L, links, ORC_HB, &rp);
    orc_rng rg = {0};
    rg.kind = ORC_RNG_GID;
    orc_sweep(&L, links, ORC_OR, &rg);
    memcpy(copy, links, sizeof(double) * nl);
    orc_halfstep_src(&L, links, copy, 1, 2, ORC_HB, &rg);
    orc_halfstep_src(&L, links, links, 2, 3, ORC_OR, &rg);
    orc_reunitarize(&L, links);

    /* measurements */
    double plq[2], poly[4], f[12];
    orc_plaquette(&L, links, plq);
    double *per_site = malloc(sizeof(double) * 2 * (size_t)V);
    if (!per_site) return 1;
    orc_plaquette_sites(&L, links, per_site);
    orc_polyakov(&L, links, poly);
    const double w = orc_wilson(&L, links, 2, 1);
    orc_fmunu(&L, links, 0, 3, sun == 3 ? 8 : 2, f);
    orc_fmunu(&L, links, 1, 1, sun == 3 ? 6 : 1, f);
    double act = 0.0;
    for (unsigned s = 0; s < V; s += 3) act += orc_link_action(&L, links, s, (int)(s % 4));
    printf("SU(%d) %dx%dx%dx%d: plaquette %.12f %.12f polyakov %.12f wilson %.12f action %.9f links %.9f\n", sun,
           n[0], n[1], n[2], n[3], plq[0], plq[1], poly[0], w, act, checksum(links, nl));
    free(per_site);
    free(hb);
    free(xt);
    free(quads);
    free(copy);
    free(links);
    return 0;
}

int main(void)
{
    /* generators on their own: RANLUX at every luxury, PM */
    float out[4 * 64 * 3];
    for (int lux = 0; lux <= 4; ++lux) {
        float *tab = malloc(sizeof(float) * 7 * 4 * 64);
        if (!tab) return 1;
        orc_ranlux_seed_table(54, 64, tab);
        orc_ranlux_produce(tab, 64, 3, orc_ranlux_nskip(lux), out);
        orc_ranlux_cpu_stream(54, orc_ranlux_nskip(lux), 100, out);
        free(tab);
    }
    uint32_t pt[4 * 64];
    orc_pm_seed_table(54, 64, pt);
    orc_pm_produce(pt, 64, 3, out);
    int32_t seed = 99;
    (void)orc_pm_one_cpu(&seed);

    const int dims[3][4] = {{6, 4, 2, 8}, {4, 6, 8, 2}, {2, 4, 6, 4}};
    for (int k = 0; k < 3; ++k) {
        if (run_lattice(3, dims[k], 5.9)) return 1;
        if (run_lattice(2, dims[k], 2.3)) return 1;
    }
    return 0;
}
