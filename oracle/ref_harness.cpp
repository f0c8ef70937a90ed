// Golden-vector harness for the reference CPU path (suncpp, CPU_RUN build).
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// reference's own sources, where they lie under the reference tree, by
// oracle/build_ref.sh into oracle/_ref/ref_harness.  It (re)generates the
// fixtures under tests/golden/ and, through its "bench" mode, is the timed
// CPU baseline of bench.py; nothing on the product path links it.
//
// What it drives (reference file:line):
//   PRNG::initialize_CPU / produce_CPU   random/random.cpp:174-213, 301-316
//   modelCPU<su_n>::lattice_initializeCPU suncpp/sunh.h:112-131
//   lattice_update_{odd,even}            suncpp/Update/sun_update.h:80-117
//   plqConf                              suncpp/Measurements/Plq.h:47-94
// and repeats the sweep order of lattice_simulateCPU (suncpp/suncpp.h:159-168):
// odd X,Y,Z,T then even X,Y,Z,T per sweep.
//
// Usage:
//   ref_harness prng <GEN> <randseries> <count> <out.bin>
//       GEN in RANLUX0..RANLUX4, PM; writes <count> float32 values.
//   ref_harness sweep <sun> <n1> <n2> <n3> <n4> <beta> <nhit> <ints> <GEN>
//                     <randseries> <nsweeps> <out_prefix>
//       ints: 1 = cold, 2 = gid.  Writes <out_prefix>_s<k>.bin (links as
//       float64, [site][dir][row][col][re,im]) after each sweep k (k=0 is the
//       start) and <out_prefix>_plq.txt (sweep, spat, temp, total).
//   ref_harness bench <sun> <n1> <n2> <n3> <n4> <beta> <nhit> <GEN>
//                     <randseries> <nsweeps> [<warmup> [<progress>]]
//       Cold start and <warmup> untimed sweeps (default 0), then times
//       <nsweeps> heatbath sweeps of suncpp on one thread; prints
//       "sweeps seconds link_updates_per_second plaquette".  With
//       <progress> = 1 a line goes to stderr after every timed half-step.

#include "suncpp/Update/sun_update.h"
#include "suncpp/Measurements/Plq.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

static PRNG_CL::PRNG::PRNG_generators gen_from_name(const char* s) {
    using P = PRNG_CL::PRNG;
    if (!strcmp(s, "RANLUX0")) return P::PRNG_generator_RANLUX0;
    if (!strcmp(s, "RANLUX1")) return P::PRNG_generator_RANLUX1;
    if (!strcmp(s, "RANLUX2")) return P::PRNG_generator_RANLUX2;
    if (!strcmp(s, "RANLUX3")) return P::PRNG_generator_RANLUX3;
    if (!strcmp(s, "RANLUX4")) return P::PRNG_generator_RANLUX4;
    if (!strcmp(s, "PM")) return P::PRNG_generator_PM;
    fprintf(stderr, "unknown generator %s\n", s);
    exit(2);
}

template <typename su_n>
static void dump(modelCPU<su_n>* m, const char* prefix, int k) {
    char name[512];
    snprintf(name, sizeof(name), "%s_s%d.bin", prefix, k);
    FILE* f = fopen(name, "wb");
    if (!f) { perror(name); exit(3); }
    fwrite(m->lattice_tableCPU, sizeof(su_n), (size_t)m->lattice_ndCPU * m->lattice_sitesCPU, f);
    fclose(f);
}

template <typename su_n>
static int run_sweeps(int argc, char** argv) {
    modelCPU<su_n> m;
    m.lattice_ndCPU = 4;
    m.lattice_size = (int*)calloc(4, sizeof(int));
    for (int i = 0; i < 4; ++i) m.lattice_size[i] = atoi(argv[3 + i]);
    m.beta = atof(argv[7]);
    m.nhit = atoi(argv[8]);
    m.ints = atoi(argv[9]);
    m.lattice_group = atoi(argv[2]);
    m.lattice_sitesCPU = 1;
    for (int i = 0; i < 4; ++i) m.lattice_sitesCPU *= m.lattice_size[i];
    PRNG_CL::PRNG prng;
    prng.PRNG_generator = gen_from_name(argv[10]);
    prng.PRNG_randseries = atoi(argv[11]);
    prng.PRNG_precision = PRNG_CL::PRNG::PRNG_precision_single;
    prng.initialize_CPU();
    int nsweeps = atoi(argv[12]);
    const char* prefix = argv[13];

    m.create_latticeCPU();
    m.lattice_initializeCPU();
    char name[512];
    snprintf(name, sizeof(name), "%s_plq.txt", prefix);
    FILE* fp = fopen(name, "w");
    for (int k = 0; k <= nsweeps; ++k) {
        if (k > 0) {
            for (int dir = 0; dir < 4; ++dir) lattice_update_odd(&m, dir, &prng);
            for (int dir = 0; dir < 4; ++dir) lattice_update_even(&m, dir, &prng);
        }
        hgpu_double total = 0.0;
        hgpu_complex p = plqConf(&m, &total);
        fprintf(fp, "%d %.17e %.17e %.17e\n", k, p.re, p.im, total);
        dump(&m, prefix, k);
    }
    fclose(fp);
    m.delete_latticeCPU();
    free(m.lattice_size);
    return 0;
}

template <typename su_n>
static int run_bench(int argc, char** argv) {
    modelCPU<su_n> m;
    m.lattice_ndCPU = 4;
    m.lattice_size = (int*)calloc(4, sizeof(int));
    for (int i = 0; i < 4; ++i) m.lattice_size[i] = atoi(argv[3 + i]);
    m.beta = atof(argv[7]);
    m.nhit = atoi(argv[8]);
    m.ints = 1;
    m.lattice_group = atoi(argv[2]);
    m.lattice_sitesCPU = 1;
    for (int i = 0; i < 4; ++i) m.lattice_sitesCPU *= m.lattice_size[i];
    PRNG_CL::PRNG prng;
    prng.PRNG_generator = gen_from_name(argv[9]);
    prng.PRNG_randseries = atoi(argv[10]);
    prng.PRNG_precision = PRNG_CL::PRNG::PRNG_precision_single;
    prng.initialize_CPU();
    int nsweeps = atoi(argv[11]);
    m.create_latticeCPU();
    m.lattice_initializeCPU();
    const int nwarm = argc >= 13 ? atoi(argv[12]) : 0;
    for (int k = 0; k < nwarm; ++k) {
        for (int dir = 0; dir < 4; ++dir) lattice_update_odd(&m, dir, &prng);
        for (int dir = 0; dir < 4; ++dir) lattice_update_even(&m, dir, &prng);
    }
    const bool progress = argc >= 14 && atoi(argv[13]) == 1;
    timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < nsweeps; ++k) {
        for (int dir = 0; dir < 4; ++dir) {
            lattice_update_odd(&m, dir, &prng);
            if (progress) fprintf(stderr, "ref_harness: sweep %d odd direction %d done\n", k, dir);
        }
        for (int dir = 0; dir < 4; ++dir) {
            lattice_update_even(&m, dir, &prng);
            if (progress) fprintf(stderr, "ref_harness: sweep %d even direction %d done\n", k, dir);
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    hgpu_double total = 0.0;
    plqConf(&m, &total);
    printf("%d %.6f %.6e %.12f\n", nsweeps, sec, 4.0 * m.lattice_sitesCPU * nsweeps / sec, total);
    m.delete_latticeCPU();
    free(m.lattice_size);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 12 && !strcmp(argv[1], "bench")) {
        int sun = atoi(argv[2]);
        if (sun == 2) return run_bench<su_2>(argc, argv);
        if (sun == 3) return run_bench<su_3>(argc, argv);
    }
    if (argc >= 6 && !strcmp(argv[1], "prng")) {
        PRNG_CL::PRNG prng;
        prng.PRNG_generator = gen_from_name(argv[2]);
        prng.PRNG_randseries = atoi(argv[3]);
        prng.PRNG_precision = PRNG_CL::PRNG::PRNG_precision_single;
        prng.initialize_CPU();
        int n = atoi(argv[4]);
        float* buf = (float*)calloc(n, sizeof(float));
        prng.produce_CPU(buf, n);
        FILE* f = fopen(argv[5], "wb");
        if (!f) { perror(argv[5]); return 3; }
        fwrite(buf, sizeof(float), n, f);
        fclose(f);
        free(buf);
        return 0;
    }
    if (argc >= 14 && !strcmp(argv[1], "sweep")) {
        int sun = atoi(argv[2]);
        if (sun == 2) return run_sweeps<su_2>(argc, argv);
        if (sun == 3) return run_sweeps<su_3>(argc, argv);
    }
    fprintf(stderr, "usage: ref_harness prng|sweep ...\n");
    return 2;
}
