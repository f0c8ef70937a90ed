/*
 * qcd_hip.h -- C ABI of libqcdhip.so, the MI355X (gfx950) implementation of
 * QCDGPU's SU(N) Monte Carlo update and measurement path.
 *
 * This header replaces the device boundary the reference host program talks
 * to: GPU_CL::GPU (SU(N)/clinterface/clinterface.h:54-415) plus the PRNG_CL::PRNG
 * produce/initialize calls (SU(N)/random/random.h:61-80) and the kernels the
 * driver binds by name in model::lattice_make_programs
 * (SU(N)/suncl/suncl.cpp:3265-3648).  Instead of an OpenCL-style "kernel by
 * name + positional arguments" registry, each reference kernel family is one
 * entry point with plain pointers and sizes; the mapping is listed per entry.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero QCD_E_* code on error;
 *    qcd_last_error() gives the message.  The reference treats every OpenCL
 *    error as fatal (clinterface.cpp:225-283); callers that want that
 *    behaviour wrap each call and exit on a non-zero return.
 *  - Calls are asynchronous on the lattice's own HIP stream unless stated;
 *    qcd_lattice_sync() waits (the reference's kernel_run waits after every
 *    launch, clinterface.cpp:1251-1278).
 *  - Host link arrays use the canonical layout
 *        links[((site*4 + dir)*N + row)*N + col][re, im]   (float64)
 *    with site = y + z*N2 + t*N2*N3 + x*N2*N3*N4 (sun_common.cl:62-66).
 */
#ifndef QCD_HIP_H
#define QCD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCD_ABI_VERSION 1

enum {
    QCD_OK = 0,
    QCD_E_ARG = 1,       /* invalid argument                                 */
    QCD_E_HIP = 2,       /* HIP runtime error                                */
    QCD_E_NOMEM = 3,     /* allocation failed                                */
    QCD_E_STATE = 4,     /* call not valid in the current state              */
    QCD_E_UNSUPPORTED = 5
};

/* Link precision (model::precision, QCDGPU.cpp:179). */
enum { QCD_FP32 = 0, QCD_FP64 = 1 };

/* Start types (model::ints: model_start_hot/cold/gid, suncl.h:67).
   QCD_START_HOT is a Haar-random start (no counterpart in the reference);
   QCD_START_RANDOM is the reference's model_start_hot itself
   (lattice_init_hot_X..T, suncl.cl:56-161), written as ints = 0 in .qcg
   headers. */
enum { QCD_START_HOT = 0, QCD_START_COLD = 1, QCD_START_GID = 2, QCD_START_RANDOM = 3 };

/* Generators (PRNG::PRNG_generators, random.h:13-27, the subset on the path). */
enum { QCD_PRNG_XOR128 = 1, QCD_PRNG_RANLUX = 2, QCD_PRNG_PM = 3 };

/* Stream semantics of the per-site generators.
 *  QCD_STREAM_REFERENCE: stream i is reference instance i with
 *    PRNG_instances = PRNGSTEP; every heatbath half-step advances each stream
 *    by exactly S = 1 + (align64(3*PRNGSTEP*(NHIT+1)*NHITPar) - 1)/PRNGSTEP
 *    quads (one PRNG::produce, random.cpp:278-284), so the numbers a link
 *    update consumes are bit-identical to the reference's PRN buffer.
 *  QCD_STREAM_FAST: same generators, each stream only advances by the quads
 *    actually consumed.                                                      */
enum { QCD_STREAM_REFERENCE = 0, QCD_STREAM_FAST = 1 };

/* Update kinds. */
enum { QCD_HEATBATH = 0, QCD_OVERRELAX = 1 };

/* Measurements (bit mask). */
enum {
    QCD_MEAS_PLAQUETTE = 1,
    QCD_MEAS_POLYAKOV = 2,
    QCD_MEAS_WILSON = 4,
    QCD_MEAS_POLYAKOV_SLICES = 8,
    QCD_MEAS_FMUNU = 16,   /* F_munu, H field: planes (x,y), (x,z), (y,z)  */
    QCD_MEAS_F0MU = 32,    /* F_munu, E field: planes (x,t), (y,t), (z,t)  */
    QCD_MEAS_ACTION_SLICES = 64   /* plaquette sums per x / y / z slice      */
};

typedef struct {
    int      sun;            /* 2 or 3 (GROUP)                                  */
    int      n[4];           /* N1..N4 = extents along x, y, z, t (L1..L4)      */
    int      precision;      /* QCD_FP32 / QCD_FP64                             */
    int      storage_reals;  /* SU(3): 18 or 12; SU(2): 4                       */
    double   beta;           /* BETA (the update uses BETA / N)                 */
    int      nhit;           /* NHIT                                            */
    int      nhitpar;        /* NHITPar (only 1 is implemented)                 */
    int      prng;           /* QCD_PRNG_*                                      */
    int      ranlux_luxury;  /* 0..4 for QCD_PRNG_RANLUX                        */
    int      stream_mode;    /* QCD_STREAM_*                                    */
    unsigned randseries;     /* RANDSERIES (0 = time based, as the reference)   */
    int      max_measurements; /* capacity of the on-device measurement arrays  */
    unsigned prng_instances; /* PRNG_instances (random.cpp:217-223): 0 = PRNGSTEP
                                = align64(V/2), one fused per-site stream per
                                link of a half-step; otherwise a multiple of 64
                                (the reference's automatic choice is the
                                device's CL_DEVICE_IMAGE3D_MAX_WIDTH,
                                clinterface.cpp:419): every heatbath half-step
                                then runs the reference's produce() of
                                1 + (align64(3*PRNGSTEP*(NHIT+1)) - 1)/instances
                                quads per instance into a PRN buffer (quad
                                sample*instances + instance) and link i reads
                                quad i + k*PRNGSTEP, for XOR128, PM and RANLUX
                                alike.  Needs QCD_STREAM_REFERENCE; no Haar
                                start and no batch with it.               */
} qcd_params;

typedef struct qcd_lattice qcd_lattice;

/* ---- device (GPU::device_select / device_initialize / device_finalize,
        clinterface.cpp:347-459) ------------------------------------------ */
int         qcd_device_open(int platform, int device);
int         qcd_device_close(void);
int         qcd_device_info(char *name, int name_len, int *compute_units, size_t *global_mem);
const char *qcd_last_error(void);
int         qcd_abi_version(void);

/* ---- lattice lifetime (model::model_lattice_init, suncl.cpp:2212-2384;
        lattice_create_buffers, :3877-3991) --------------------------------- */
int  qcd_lattice_create(const qcd_params *params, qcd_lattice **out);
int  qcd_lattice_destroy(qcd_lattice *lat);
int  qcd_lattice_sync(qcd_lattice *lat);
/* Launch-shape knobs of the update kernels.  They only change scheduling,
   never results (the 64-link units of a half-step are independent):
   QCD_TUNE_XCD_MAP       which 64-link units each of the 8 XCDs (hardware
                          workgroup b runs on XCD b mod 8) updates:
                          0: hardware order;
                          1: one contiguous range of compact site indices
                          per XCD, an x-slab;
                          2 (default): the same x-slabs, dispatched
                          cross-section first (x offset, then t, then z)
                          for L2 reuse;
                          3: each XCD owns a block of x-values and z-rows,
                          N1/ax x N3/az with ax*az = 8, chosen for the
                          fewest halo links (1/x-extent + 1/z-extent),
                          dispatched cross-section first (slower than 2 on
                          MI355X at 32^3 x 8: 0.763 vs 0.744 ms per sweep).
                          Modes 2 and 3 need N2/2 to divide 64 and whole
                          z-groups and units per XCD; otherwise mode 1 runs.
   QCD_TUNE_BLOCKS_PER_CU cap on resident update workgroups per CU (reserved
                          LDS), 0 = no cap (default).
   QCD_TUNE_UPDATE_SHAPE  3: three wavefronts share 64 links, one staple
                          direction each (default); 1: one wavefront per 64
                          links, workgroups of four.
   QCD_TUNE_SHADOW        race check (new; the reference has none): 1 makes
                          every update half-step gather its staples from a
                          copy of the table taken just before it (one more
                          table-sized device buffer and one device copy per
                          half-step) and write into the table, 0 (default)
                          updates in place.  A half-step reads only links it
                          does not write, so the two give identical bits; a
                          difference means a kernel read a link that the same
                          launch writes.  Single lattices only.
   QCD_TUNE_CYCLES_PER_GRAPH  K (1..64, default 1): qcd_cycle_capture also
                          records K consecutive cycles as one graph in which
                          each cycle's measurement (plaquette fused into the
                          last sweep, Polyakov loop, final sums) runs on a
                          second stream beside the next cycle's odd X, Y, Z
                          half-steps, which touch nothing it reads or
                          writes; the odd T half-step waits for it.
                          qcd_cycle_launch / _time replay whole K-blocks,
                          then single cycles.  Only for masks of the
                          plaquette and the Polyakov loop with the plaquette
                          fused (otherwise single cycles).  Results are the
                          same bits; qcd_lattice_get_tune's detail[0] gives
                          the K of the captured graph (0: none).  Measured
                          on MI355X at 32^3 x 8: K = 8 is 1.7 % slower than
                          single cycles (the overlapped launches slow the
                          half-steps more than the measurement costs).     */
enum { QCD_TUNE_XCD_MAP = 1, QCD_TUNE_BLOCKS_PER_CU = 2, QCD_TUNE_UPDATE_SHAPE = 3, QCD_TUNE_SHADOW = 4,
       QCD_TUNE_CYCLES_PER_GRAPH = 5 };
int  qcd_lattice_tune(qcd_lattice *lat, int knob, int value);
/* ---- kernels by name (GPU::kernel_init / kernel_run, clinterface.cpp:1132-1278,
        with the names lattice_make_programs creates, suncl.cpp:3265-3648,
        and the PRNG's, random.cpp:487-493,668,886), for a host written
        against the reference's registry: update_{odd,even}_{X,Y,Z,T} (a
        heatbath half-step, its PRN production included),
        overrelax_{odd,even}_{X,Y,Z,T} (new), lattice_init_cold / _gid,
        lattice_init_hot_X (the reference's whole hot start; _Y, _Z, _T run
        nothing), lattice_GramSchmidt, lattice_measurement_plq,
        lattice_measurement (with F_munu), lattice_polyakov,
        lattice_measurement_wilson, lattice_polyakov_diff_x and
        lattice_action_diff_x (the slices of all three directions; _y, _z
        nothing); the reduce_* kernels, clear_measurement and the PRNG
        kernels xor128 / pm / ranlux are accepted and run nothing (their
        work is part of the launches above).  Ids count from 1.  _run
        enqueues and waits (as kernel_run); _enqueue only enqueues.         */
int  qcd_kernel_create(qcd_lattice *lat, const char *name, int *id);
int  qcd_kernel_run(qcd_lattice *lat, int id);
int  qcd_kernel_enqueue(qcd_lattice *lat, int id);
/* Name and number of runs (kernel_number_of_starts) of kernel `id`.        */
int  qcd_kernel_info(const qcd_lattice *lat, int id, char *name, int name_len, unsigned *starts);
/* Records the kernels enqueued between _begin and _end as the lattice's
   graph (replacing a captured cycle), replayed by qcd_cycle_launch /
   qcd_cycle_time.  With fused XOR128 / PM streams the recorded heatbath
   half-steps must be even in number.                                       */
int  qcd_graph_begin(qcd_lattice *lat);
int  qcd_graph_end(qcd_lattice *lat);
/* ---- device buffers by id (GPU::buffer_init, buffer_write, buffer_map,
        buffer_kill, buffer_size_align, clinterface.cpp:1296-1440,1558-1590),
        for a host that keeps its own buffers in the reference's registry
        style.  Types as GPU_buffer_types (clinterface.h:73-82): INPUT, IO and
        CONSTANT copy `count * elem_size` bytes from host_ptr at creation,
        GLOBAL, OUTPUT and UAV start zeroed, LDS records a size and holds no
        memory.  _write uploads the whole host copy again (blocking);
        _map downloads the buffer into a host mirror the library owns until
        _kill and returns it (blocking, as buffer_map with CL_TRUE); _unmap
        writes the mirror back.  Ids count from 1.  _size_align rounds `size`
        up to a multiple of `step` (0: the device's largest workgroup,
        1024).  The lattice's own data in the reference's layouts goes
        through qcd_lattice_export / qcd_read_measurements / qcd_qcg_*.      */
enum { QCD_BUFFER_GLOBAL = 1, QCD_BUFFER_INPUT = 2, QCD_BUFFER_IO = 3, QCD_BUFFER_OUTPUT = 4,
       QCD_BUFFER_CONSTANT = 5, QCD_BUFFER_LDS = 6, QCD_BUFFER_UAV = 7 };
int  qcd_buffer_create(int type, size_t count, size_t elem_size, void *host_ptr, int *id);
int  qcd_buffer_write(int id);
int  qcd_buffer_map(int id, void **host);
int  qcd_buffer_unmap(int id);
int  qcd_buffer_info(int id, size_t *bytes, void **device_ptr);
int  qcd_buffer_kill(int id);
unsigned qcd_buffer_size_align(unsigned size, unsigned step);
/* ---- kernel timing (GPU::print_time_detailed, clinterface.cpp:1626-1666,
        shown when GPU_debug.profiling is set, suncl.cpp:265).  With it on,
        every kernel family the lattice launches eagerly (update half-steps,
        PRN production, starts, measurement kernels, final sums) is timed
        with events on the lattice stream and waited for; a captured cycle
        is refused while it is on.  Turning it on clears the table.         */
int  qcd_lattice_profile(qcd_lattice *lat, int on);
int  qcd_profile_count(const qcd_lattice *lat, int *n);
/* Entry i: the family's name, launches, mean and standard deviation in ms. */
int  qcd_profile_entry(const qcd_lattice *lat, int i, char *name, int name_len, unsigned *starts, double *mean_ms,
                       double *dev_ms);
/* ---- x-slab: one GPU's share of a lattice cut into slabs along x, as the
        reference's multi-device BIGLAT path holds it (suncl.cpp:2422: every
        device keeps Nx + 2 x-slices; UpdateEdges, suncl.cpp:4005-4062,
        refreshes the two edge slices).  The lattice's N1 counts the two halo
        slices x = 0 and x = N1 - 1: with the slab on, half-steps update the
        interior slices 1 .. N1 - 2 only, reading the halos as they stand (no
        periodic wrap along x is ever taken); measurements still cover the
        whole table, and cycles measure the plaquette with its own kernel.
        Needs N1 >= 4 and N2 N3 N4 / 2 divisible by 64; no twisted
        boundary.  Parities are the slab's own (x counted from its first
        halo slice).                                                        */
int  qcd_lattice_set_slab(qcd_lattice *lat, int on);
/* Bytes of one face: the links of one parity and direction in an x-slice. */
int  qcd_slab_face_bytes(const qcd_lattice *lat, size_t *bytes);
/* Device-to-device copy, on the lattice's stream, of the (parity, mu) links
   of one face: unpack = 0 copies the first (side 0) or last (side 1)
   interior slice into dev_buf, unpack = 1 copies dev_buf into the halo on
   that side.  After each half-step (parity, mu), a slab sends both of its
   boundary faces of that group and receives both halos -- what a multi-GPU
   run moves with RCCL send / recv between x-neighbours.                    */
int  qcd_slab_face(qcd_lattice *lat, int side, int parity, int mu, void *dev_buf, int unpack);
/* The value a knob takes in the update launches of this lattice: for
   QCD_TUNE_XCD_MAP the mode that applies after the fallbacks above, and in
   detail[0..1] (optional) the XCD block's x-values and z-rows (0 for modes
   0 and 1).  No counterpart in the reference (its work-group order is the
   OpenCL runtime's).                                                        */
int  qcd_lattice_get_tune(const qcd_lattice *lat, int knob, int *value, int *detail);
/* Bytes of device memory held by the link table. */
int  qcd_lattice_table_bytes(const qcd_lattice *lat, size_t *bytes);

/* ---- starts: kernels lattice_init_cold / lattice_init_gid (suncl.cl:164-347);
        QCD_START_HOT draws Haar-distributed links from the per-site streams
        (the reference's lattice_init_hot_* is not Haar, su2_update_cl.cl:110),
        so it needs qcd_prng_init first; stream i fills the links of the two
        sites with compact index i (parity 0 then 1, directions 0..3).
        QCD_START_RANDOM is the reference's hot start
        (suncl.cpp:5234-5248): it seeds the streams itself (qcd_prng_init
        is not needed), sizes every produce() of the run as the reference's
        hot runs do (suncl.cpp:2304-2305: S_hot = ceil(align64(4 ROWSIZE +
        3 PRNGSTEP (NHIT+1)) / PRNGSTEP) quads per stream, ROWSIZE =
        align64(V)), and fills direction d from the (d+2)-th produce():
        site s (lexicographic index y + N2 (z + N3 (t + N4 x))) takes M. Di
        Pierro's random SU(2) element of quad s (lattice_random2,
        su2_update_cl.cl:110-146), or for SU(3) the product E12(A) E13(B)
        E23(C) of the elements of components x, y, z of quads s,
        s + ROWSIZE, s + 2 ROWSIZE (lattice_random3,
        su3_update_cl.cl:165-225).  Counter 5 afterwards.  Needs
        QCD_STREAM_REFERENCE (RANLUX always is).  The other starts return
        the lattice to the cold runs' produce() size (re-seeding streams
        seeded under the other size); a change of size drops a captured
        cycle.                                                             */
int  qcd_lattice_init(qcd_lattice *lat, int start);

/* ---- PRNG (PRNG::initialize, random.cpp:216-277): seeds the per-site
        streams through glibc srand(RANDSERIES)/rand() in the reference order
        and performs the initial produce().                                  */
int  qcd_prng_init(qcd_lattice *lat);

/* ---- host copies (buffer_init / buffer_map, clinterface.cpp:1316-1433) */
int  qcd_lattice_import(qcd_lattice *lat, const double *links);
int  qcd_lattice_export(qcd_lattice *lat, double *links);   /* synchronous */

/* ---- updates: update_{even,odd}_{X,Y,Z,T} (suncl.cl:399-975) with the PRNG
        produce fused in; parity 0 = even, 1 = odd; mu = 0..3 (X,Y,Z,T).     */
int  qcd_halfstep(qcd_lattice *lat, int parity, int mu, int kind);
/* One sweep = odd X,Y,Z,T then even X,Y,Z,T (suncl.cpp:5392-5409). */
int  qcd_sweep(qcd_lattice *lat, int kind);
/* Whole-lattice reunitarisation: lattice_GramSchmidt (suncl.cl:350-396),
   in FP64 for both link precisions.                                         */
int  qcd_reunitarize(qcd_lattice *lat);

/* ---- measurements: lattice_measurement_plq + reduce_measurement_plq_double2
        (sun_measurements_cl.cl:1178-1323), lattice_polyakov +
        reduce_polyakov_double2 (polyakov.cl:54-431).  Each call appends one
        record on the device; sums are FP64 and unnormalised:
          plaquette: (sum ReTr spatial, sum ReTr temporal)
          polyakov : (sum ReTr P, sum ImTr P, sum |Tr P|^2, sum |Tr P|^4)   */
int  qcd_measure(qcd_lattice *lat, int mask);
/* qcd_measure without allocating: no reference counterpart.  For callers
   that cannot allocate device memory (stream capture, serving loops): the
   partial sums and per-site buffers of every measurement kernel live in one
   fixed 16 MiB device array per GPU instead of the lattice's buffers or the
   action slices' buffer allocated on first use.  Lattices whose buffers for
   `mask` exceed it fail with QCD_E_ARG before any launch (SU(3) 32^3 x 8
   with every measurement fits).  Records are those qcd_measure appends, bit
   for bit.  Calls that overlap on one GPU share the array and race.      */
int  qcd_measure_noalloc(qcd_lattice *lat, int mask);
/* Copies the records so far (synchronous).  plq: 2*count, poly: 4*count.  */
int  qcd_read_measurements(qcd_lattice *lat, double *plq, double *poly, int *count_plq, int *count_poly);
/* ---- Wilson loop: lattice_measurement_wilson + reduce (wilson_loop.cl:55-465;
        CPU twin su3cpu.cpp:1435-1522), R links along x, y and z, T along t
        (wilson_R / wilson_T, QCDGPU.cpp:202-203, defaults 7 and 2).  Each
        QCD_MEAS_WILSON measurement appends sum ReTr W over sites and the
        three spatial directions (the reference divides by 3V).            */
int  qcd_lattice_set_wilson(qcd_lattice *lat, int R, int T);
/* ---- twisted boundary conditions (PHI / OMEGA, the reference's -D TBC:
        lattice_table_3 / lattice_table_2, su3_matrix_memory.cl:57-98,
        su2_matrix_memory.cl:56-88; PHI / OMEGA in lattice_parameters[1..2],
        suncl.cpp:3941-3951).  Every read of a Y link of a site with
        x = N1 - 1 -- in the updates (the link itself and the staples), the
        plaquette, the Wilson loop and F_munu -- sees T U, with
        T = diag(e^{i(phi+omega)/2}, e^{i(omega-phi)/2}, e^{-i omega}) for
        SU(3) and diag(e^{i phi/2}, e^{-i phi/2}) for SU(2) (omega unused);
        an update stores its result as computed, without removing T, as the
        reference does.  phi = omega = 0 (the default) turns the twist off.
        Set it before qcd_cycle_capture (a captured cycle keeps the values
        it was captured with); while it is on, cycles measure the plaquette
        with the standalone kernel.  Export, import and the state image hold
        the stored links.                                                    */
int  qcd_lattice_set_twist(qcd_lattice *lat, double phi, double omega);
int  qcd_lattice_get_twist(qcd_lattice *lat, double *phi, double *omega);
int  qcd_read_wilson(qcd_lattice *lat, double *w, int *count);
/* ---- Polyakov loop per slice (PL_level > 2: lattice_polyakov_diff_x/y/z +
        reduce_polyakov_diff_*_double2, polyakov.cl:142-560): a
        QCD_MEAS_POLYAKOV_SLICES measurement (it implies QCD_MEAS_POLYAKOV)
        appends, for every slice x = 0..N1-1, y = 0..N2-1, z = 0..N3-1, the
        sums (sum Re Tr P, sum Im Tr P) over the slice's spatial sites.
        out[(k*(N1+N2+N3) + slice)*2 + c], slices in the order x, y, z.      */
int  qcd_read_polyakov_slices(qcd_lattice *lat, double *out, int *count);
/* ---- action per slice (get_actions_diff: lattice_action_diff_x/y/z +
        reduce_action_diff_*_double2, sun_measurements_cl.cl:186-708; the
        reference's QCDGPU.cpp:218 sets get_actions_diff = false and no
        init.dat key turns it on).  A QCD_MEAS_ACTION_SLICES measurement
        (it implies QCD_MEAS_PLAQUETTE, measured with the standalone kernel)
        stores, beside that plaquette record, for every slice x = 0..N1-1,
        y = 0..N2-1, z = 0..N3-1 the sums over the slice's sites of the
        three spatial and of the three temporal ReTr U_p:
        out[(rec*(N1+N2+N3) + slice)*2 + (0 spatial, 1 temporal)]; records
        of plaquette measurements without it hold zeros.  The reference's
        S per slice is beta/N (N - sum / (3 * sites in the slice)).  The
        first such measurement allocates 16 bytes per site.                  */
int  qcd_read_action_slices(qcd_lattice *lat, double *out, int *count);
/* ---- field strength (get_Fmunu / get_F0mu: lattice_measurement_plq with
        FMUNU / F0MU, sun_measurements_cl.cl:710-1173, trace_lambda* and
        trace_sigma*, su3_measurements_cl.cl:357-460,
        su2_measurements_cl.cl:89-123).  A QCD_MEAS_FMUNU or QCD_MEAS_F0MU
        measurement (one of the two; either implies QCD_MEAS_PLAQUETTE)
        stores, beside that plaquette record, for the two generators and the
        three planes, sum over sites of F = -i Tr(P G) / 2 with
        P = U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+ in FP64:
        out[rec*12 + gen*6 + plane*2 + (0 re, 1 im)].  Records of plaquette
        measurements taken without F hold zeros.  SU(2) imaginary parts are 0
        (the reference's trace_sigma*).                                       */
/* Generators (Fmunu_index1/2, suncl.cpp:2093-2113; init keys FMUNU1/2/4 and
   FMUNU5/6/7): SU(3) gen1 in 1..4, gen2 in 5..8 (Gell-Mann lambda_k, default
   3 and 8); SU(2) gen1 1 or 3, gen2 2 (Pauli sigma_k, default 3 and 2).   */
int  qcd_lattice_set_fmunu(qcd_lattice *lat, int gen1, int gen2);
int  qcd_read_fmunu(qcd_lattice *lat, double *out, int *count);
int  qcd_reset_measurements(qcd_lattice *lat);

/* ---- graphs: one captured cycle of n_hb heatbath sweeps, n_or
        overrelaxation sweeps and the measurements in `mask`, replayed with
        qcd_cycle_launch.  Replaces the synchronous per-kernel loop of
        model::lattice_simulate (suncl.cpp:5391-5483).                       */
int  qcd_cycle_capture(qcd_lattice *lat, int n_hb, int n_or, int mask);
/* The same cycle enqueued eagerly (no graph).  In both forms the plaquette
   of the configuration left by the last sweep is accumulated inside that
   sweep's even Y, Z and T half-steps (each plaquette of those planes holds
   exactly one even-site link of the later direction), so no separate
   plaquette kernel runs; the Polyakov loop is its own kernel.              */
int  qcd_cycle_run(qcd_lattice *lat, int n_hb, int n_or, int mask);
int  qcd_cycle_launch(qcd_lattice *lat, int repeats);
/* Wall time of `repeats` graph launches, bracketed by HIP events on the
   lattice stream (synchronous).                                             */
int  qcd_cycle_time(qcd_lattice *lat, int repeats, double *ms);
/* Runs `repeats` cycles eagerly with HIP events around every update kernel
   and reports the mean kernel duration per kind (us) and the total (ms).   */
int  qcd_cycle_profile(qcd_lattice *lat, int repeats, double *hb_us, double *or_us,
                       double *meas_us, double *total_ms);

/* ---- checkpoints (model::lattice_save_state / lattice_load_state /
        lattice_make_bin_header / lattice_load_bin_header,
        suncl.cpp:1857-2079; generator replay suncl.cpp:5349-5351).        */
int  qcd_lattice_get_params(const qcd_lattice *lat, qcd_params *out);
/* PRNG::produce calls the reference would have made so far (its
   PRNG_counter, random.cpp:283): 1 after qcd_prng_init, +4 for a hot start
   (one produce before each of the four init kernels, suncl.cpp:5236-5248),
   +1 per heatbath half-step.  Not kept for batch chains.                   */
int  qcd_prng_counter(const qcd_lattice *lat, unsigned *counter);
/* Re-seeds the streams (qcd_prng_init) and advances them to production
   `counter` of a run that began with `start` (QCD_START_*), as the
   reference's resume does (while (PRNG_counter > PRNG0->PRNG_counter)
   produce()), with the produce() size of that start (QCD_START_RANDOM:
   the hot runs').  XOR128 / PM heatbath productions are applied as one jump.
   Fast-layout XOR128 / PM streams cannot be replayed (QCD_E_UNSUPPORTED).  */
int  qcd_prng_replay(qcd_lattice *lat, unsigned counter, int start);
/* Exact state image: the table as stored on the device, the current
   generator states, all records and their counts, the production counter
   and the start flags, behind a head with the lattice parameters.  Loading
   it into a lattice created with equal qcd_params continues the run
   bit-identically (any other lattice: QCD_E_ARG).  Synchronous.            */
int  qcd_lattice_state_bytes(const qcd_lattice *lat, size_t *bytes);
int  qcd_lattice_state_save(qcd_lattice *lat, void *buf, size_t bytes);
int  qcd_lattice_state_load(qcd_lattice *lat, const void *buf, size_t bytes);
/* Replaces the records and their counts (the reference's buffer_init of the
   measurement arrays from the loaded host copies, suncl.cpp:2061-2068);
   counts beyond max_measurements are kept as counts, the extra records are
   dropped as qcd_measure drops them.  plq: 2*count, poly: 4*count.         */
int  qcd_write_measurements(qcd_lattice *lat, const double *plq, const double *poly, const double *wil,
                            int count_plq, int count_poly, int count_wil);

/* The reference's .qcg state file.  64 uint32 header words
   (lattice_make_bin_header, suncl.cpp:1939-1993), then, as
   lattice_save_state writes them (suncl.cpp:1889-1901): the measurement
   partials (double2 x measurement_size_F; written as zeros), the action
   sums (double2 x energies_size), the plaquette sums (double2 x
   energies_size_F), the Wilson loop sums (double x energies_size, if
   get_wilson), the Polyakov sums (Re, Im) (double2 x polyakov_loop_size,
   if pl_level > 0) and the table in the reference layout (float4 / double4
   x table_size: plane k of direction d at [gindex + (4k + d) * ROWSIZE];
   SU(3) planes (Re u11, Re u12, Re u13, Re u23), (Im u11, Im u12, Im u13,
   Im u23), (Re u21, Re u22, Im u21, Im u22); SU(2) (Re u11, Re u12, Im u11,
   Im u12)).  Sizes use buffer_size_align with the reference driver's
   64-item work-group cap (QCDGPU.cpp:74).  After the table this library
   appends "QCDGPUX1", a uint64 byte count and the exact state image; the
   reference stops reading before it.  With F_munu (get_fmunu or get_f0mu)
   the partials and plaquette sections are 7 times longer (suncl.cpp:
   2266-2298) and the plaquette section holds, after the plaquette sums,
   the F sums of the three planes for the first generator, then for the
   second (double2 (re, im) per record; reduce_measurement_plq_double2,
   sun_measurements_cl.cl:1312-1354).                                      */
typedef struct {
    unsigned init;            /* INIT (0 = continue a run)            word 4     */
    unsigned start;           /* ints: QCD_START_*                    word 5     */
    unsigned prng_counter;    /* production counter                   word 9
                                 (the writer takes it from the lattice)         */
    unsigned nav, nav_counter, niter, iter, iter_counter;   /* words 10-14      */
    unsigned wilson_r, wilson_t;                            /* words 16-17      */
    unsigned pl_level;                                      /* word 25          */
    unsigned get_actions, get_plaquettes, get_wilson;       /* words 29-31      */
    unsigned fmunu_index1, fmunu_index2, get_fmunu;         /* words 26-28      */
    unsigned get_f0mu;        /* E-field F_munu: not a header word; a file
                                 with F sections (energies_size_F = 7 x
                                 energies_size, word 34) and get_fmunu 0       */
    double   phi, omega;      /* PHI, OMEGA (words 21-24; the reader reports
                                 them and applies them to the lattice, the
                                 writer takes them from the lattice)           */
} qcd_run_info;
/* Writes the state file (atomically: a temporary file renamed into place).  */
int  qcd_qcg_write(qcd_lattice *lat, const char *path, const qcd_run_info *info);
/* Reads the header: the lattice parameters (storage, stream layout and
   record capacity from the exact state image when the file has one, else
   18 / 4 reals, reference streams and ITER records) and the run counters.
   has_exact (optional) tells whether the image is present.  No device use. */
int  qcd_qcg_read_header(const char *path, qcd_params *params, qcd_run_info *info, int *has_exact);
/* Loads the file into `lat`: from the exact image when its parameters equal
   the lattice's, else from the reference sections (table, plaquette,
   Polyakov (Re, Im) and Wilson sums; P2 / P4 are not in them) with the
   streams replayed to the header's production counter.                     */
int  qcd_qcg_read(qcd_lattice *lat, const char *path, qcd_run_info *info);

/* ---- batched ensembles (new; the reference runs one lattice per process and
        parallelises ensembles trivially, paper p.7 sec. 5).  Many independent
        chains with their own extents, beta and random series are updated by
        one launch per half-step and bucket (chains of equal spatial extent);
        each chain is a full lattice with its own records.  `common` gives
        everything else (group, precision, storage, NHIT, generator, stream
        layout, record capacity); its n, beta and randseries are ignored.     */
typedef struct {
    int      n[4];        /* N1..N4 of this chain                             */
    double   beta;        /* BETA of this chain                               */
    unsigned randseries;  /* RANDSERIES of this chain's streams               */
} qcd_chain_desc;
typedef struct qcd_batch qcd_batch;
int  qcd_batch_create(const qcd_params *common, int nchains, const qcd_chain_desc *chains, qcd_batch **out);
int  qcd_batch_destroy(qcd_batch *batch);
int  qcd_batch_info(qcd_batch *batch, int *nchains, int *nbuckets);
/* Starts every chain (QCD_START_HOT or QCD_START_COLD) and seeds its streams. */
int  qcd_batch_init(qcd_batch *batch, int start);
int  qcd_batch_sweep(qcd_batch *batch, int kind);
/* n_hb heatbath + n_or overrelaxation sweeps of every chain, then the
   measurements in `mask` (plaquette fused into the last sweep, Polyakov),
   appended to each chain's records; eagerly, or captured once and replayed. */
int  qcd_batch_cycle_run(qcd_batch *batch, int n_hb, int n_or, int mask);
int  qcd_batch_cycle_capture(qcd_batch *batch, int n_hb, int n_or, int mask);
int  qcd_batch_cycle_launch(qcd_batch *batch, int repeats);
int  qcd_batch_cycle_time(qcd_batch *batch, int repeats, double *ms);
int  qcd_batch_sync(qcd_batch *batch);
/* Autotuner: times every bucket's half-steps with each update-kernel shape
   (three wavefronts per 64 links or one) and each placement of the 64-link
   units over the 8 XCDs (0: every chain whole on one XCD, units in index
   order; 1: the same in cross-section order, x fastest, then t, then z;
   2: every chain cut into eight slabs, one per XCD) and keeps the fastest
   per bucket and update kind; chains are saved and restored around the
   timing, so results do not depend on it.  shapes_out (optional): 2 ints
   per bucket (heatbath, overrelaxation shape).  Capture the cycle after
   tuning.                                                                 */
int  qcd_batch_autotune(qcd_batch *batch, int repeats, int *shapes_out);
/* The choices for bucket `bucket`: out[0..5] = heatbath shape,
   overrelaxation shape, heatbath placement, overrelaxation placement, the
   bucket's N1 (spatial extent), chains in the bucket.                      */
int  qcd_batch_get_tuning(qcd_batch *batch, int bucket, int *out);
/* Sets bucket `bucket`'s choices by hand (shapes 1 or 3, placements 0..2). */
int  qcd_batch_set_tuning(qcd_batch *batch, int bucket, int shape_hb, int shape_or, int place_hb, int place_or);
/* Launch grouping of the batched half-steps (new): 1 (default) updates every
   chain with one launch per (parity, mu) and kernel shape, over the
   buckets' chosen work maps concatenated; 0 launches once per bucket.  The
   chains' results are the same either way.  A captured cycle keeps the
   grouping it was captured with. */
int  qcd_batch_set_merge(qcd_batch *batch, int on);
/* Chain `index` as a lattice handle (export, records, measure); the batch's
   work is complete when this returns.  The batch owns it: destroying it,
   re-sizing its streams (a reference hot start) or resuming a hot run on it
   fail with QCD_E_STATE, and so does the next batch sweep or cycle after
   the chain was swept on its own (its stream buffers would no longer be
   the batch's).                                                           */
int  qcd_batch_chain(qcd_batch *batch, int index, qcd_lattice **lat);

/* ---- learned global proposal for SU(2) (new; no reference counterpart): a
        gauge-equivariant flow of masked coupling layers on 8-channel
        plaquette / staple features (two 3^4-stencil convolutions, 32 hidden
        channels, tanh; BF16 MFMA with FP32 accumulation) and rational-
        quadratic spline maps of the active loop's angle, for an independence
        Metropolis step with the exact Wilson action; the plain heatbath stays
        the fallback.  Links are float quaternions (a0, a1, a2, a3),
        U = a0 + i a.sigma, in links[((b*4 + mu)*V + site)*4 + c].  log q is
        the density with respect to normalised Haar measure; action is
        beta * sum_P (1 - Tr U_P / 2).  Parameters: per layer W1[32][81*8],
        b1[32], W2[16][81*32], b2[16] (FP32; stencil point st = ((ox*3+oy)*3
        +oz)*3+ot with offsets o-1 along x,y,z,t, k = st*cin + channel).     */
typedef struct {
    int n[4];        /* lattice extents (site count divisible by 16)           */
    int layers;      /* coupling layers (layer l: mu = l%4, mask (l/4)%2)       */
    int hidden;      /* hidden channels: 32                                     */
    int bins;        /* spline bins, 1..5                                       */
    int max_batch;   /* configurations per sample / evaluate call               */
} qcd_flow_desc;
typedef struct qcd_flow qcd_flow;
int  qcd_flow_create(const qcd_flow_desc *desc, qcd_flow **out);
int  qcd_flow_destroy(qcd_flow *flow);
int  qcd_flow_param_count(const qcd_flow *flow, size_t *count);
int  qcd_flow_set_params(qcd_flow *flow, const float *params, size_t count);
/* Draws `batch` Haar configurations (seeded counter hash), maps them through
   the flow, returns the links, log q and the action (synchronous).        */
int  qcd_flow_sample(qcd_flow *flow, unsigned long long seed, int batch, double beta, float *links,
                     double *logq, double *action);
/* log q and the action of given configurations (the inverse flow).         */
int  qcd_flow_evaluate(qcd_flow *flow, const float *links, int batch, double beta, double *logq, double *action);
/* One learned Metropolis sweep of an SU(2) FP32 lattice of the flow's
   extents, on the lattice's stream: for every coupling layer in turn, its
   spline parameters from the current (frozen) links, then every active link
   independently replaced by the layer's map of a Haar draw with probability
   min(1, exp(-dS) q(U)/q(U')) -- exact Metropolis-within-Gibbs, since the
   layer's conditioners see frozen links only and no plaquette holds two
   active links.  Eight layers visit every link once.  Asynchronous.        */
int  qcd_flow_imh_sweep(qcd_flow *flow, qcd_lattice *lat, unsigned long long seed, double beta);
/* The sweep's seed lives on the device: qcd_flow_imh_sweep sets it;
   qcd_flow_imh_sweep_next runs the sweep with the previous seed + 1, so a
   graph recorded around it (qcd_graph_begin / _end on the lattice) replays
   sweeps with successive seeds, exactly the sweeps qcd_flow_imh_sweep would
   run with those seeds; qcd_flow_imh_seed sets the seed without a sweep
   (synchronous). */
int  qcd_flow_imh_sweep_next(qcd_flow *flow, qcd_lattice *lat, double beta);
int  qcd_flow_imh_seed(qcd_flow *flow, unsigned long long seed);
/* Training by hand-written kernels (BF16 MFMA, FP32 accumulation): the
   layer-local forward KL of `batch` heatbath configurations (host links in
   the sample / evaluate layout) -- for every coupling layer, the sum over
   its active links of -log q_a(U_a | frozen links) on the data itself, the
   density the learned Metropolis sweep proposes with -- and its gradient:
   the spline part by central differences in double precision, then conv2,
   tanh and conv1 backwards (weight gradients as MFMA GEMMs over the sites,
   the hidden gradient as the transposed stencil convolution).
   qcd_flow_train_grad returns the loss per configuration and the gradient
   summed over the batch (grad: param_count floats, optional);
   qcd_flow_train_step takes one Adam step (beta 0.9 / 0.999) on the batch
   mean and updates the flow's parameters; qcd_flow_get_params reads them. */
int  qcd_flow_train_grad(qcd_flow *flow, const float *links, int batch, double *loss, float *grad);
int  qcd_flow_train_step(qcd_flow *flow, const float *links, int batch, float lr, double *loss);
int  qcd_flow_get_params(qcd_flow *flow, float *params, size_t count);
/* Links accepted and proposed by qcd_flow_imh_sweep so far (synchronous);
   reset != 0 starts the counts again.                                       */
int  qcd_flow_imh_counts(qcd_flow *flow, unsigned long long *accepted, unsigned long long *proposed, int reset);

/* ---- reference PRN buffer semantics (PRNG::produce, random.cpp:278-284;
        kernels xor128 prngcl_xor128.cl:100, pm prngcl_pm.cl:86,
        ranlux prngcl_ranlux.cl:534).  Host seed tables in, quads out:
        quads[(i*n + g)*4 + c] for instance g, sample i.  Tables advance.
        XOR128 / PM tables: uint32 [n][4].  RANLUX: float [7][n][4] (six
        float4 of lagged values, then (23, 9, block quad counter, carry));
        ranlux_luxury 0..4 selects the skip (24, 48, 97, 223, 389).  RANLUX
        lattice streams always follow the reference layout: one produce()
        of S quads per heatbath half-step (qcd_params.stream_mode is not
        used for RANLUX).                                                     */
int  qcd_prng_produce_host(int prng, int ranlux_luxury, unsigned n_instances, unsigned samples,
                           void *table, float *quads);


#ifdef __cplusplus
}
#endif
#endif
