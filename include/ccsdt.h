/*
 * ccsdt.h -- C ABI of libccsdt.so, the MI355X (gfx950) CCSD(T) triples body.
 *
 * Two kinds of entry point:
 *
 *  1. Drop-ins with the reference's exact argument lists, column-major layouts
 *     and 1-based index conventions.  They are synchronous: the results are in
 *     the caller's host arrays on return.  Like the reference they return void;
 *     failures are reported through ccsdt_last_error() and, under the default
 *     policy, abort the process with a message (see ccsdt_set_error_policy).
 *
 *  2. An explicit-residency job API: the caller uploads a job's arrays once,
 *     then runs lists of (a,i,j,k,klo) triples on the device.  These return a
 *     status (0 = success, negative = CCSDT_E_*).
 *
 * All array arguments are double precision (IEEE binary64), all index and size
 * arguments 32-bit int, as in the reference (LP64 Fortran INTEGER).
 */
#ifndef CCSDT_H
#define CCSDT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- drop-ins */

/* One (a,i,j,k) triple: the 16 DGEMMs into f1n..f4t (overwritten; each
 * nvir*nvir, column major), eaijk, the energy reduction, and
 *   *emp4 += emp4i; *emp5 += emp5i; if (*i != *k) { *emp4 += emp4k; *emp5 += emp5k; }
 * Replaces ccsd_trpdrv_omp_cbody_, omp/ccsd_trpdrv_omp_cbody.c:31-44 (caller
 * prototype omp/test.c:39-52).  k-indexed operands (tkj, kkj, jkj, jka, kka,
 * tka, xka) are read at slice (k-klo); tia, xia, jij, kij are read over their
 * first nocc*nvir elements; dint and t1v vectors over nvir elements. */
void ccsd_trpdrv_omp_cbody_(double *f1n, double *f1t, double *f2n, double *f2t,
                            double *f3n, double *f3t, double *f4n, double *f4t,
                            double *eorb, int *ncor, int *nocc, int *nvir,
                            double *emp4, double *emp5,
                            int *a, int *i, int *j, int *k, int *klo,
                            double *tij, double *tkj, double *tia, double *tka,
                            double *xia, double *xka, double *jia, double *jka,
                            double *kia, double *kka, double *jij, double *jkj,
                            double *kij, double *kkj,
                            double *dintc1, double *dintx1, double *t1v1,
                            double *dintc2, double *dintx2, double *t1v2);

/* Same contract under the name of the OpenCL variant, ocl/ccsd_trpdrv_cbody.c:31. */
void ccsd_trpdrv_cbody_(double *f1n, double *f1t, double *f2n, double *f2t,
                        double *f3n, double *f3t, double *f4n, double *f4t,
                        double *eorb, int *ncor, int *nocc, int *nvir,
                        double *emp4, double *emp5,
                        int *a, int *i, int *j, int *k, int *klo,
                        double *tij, double *tkj, double *tia, double *tka,
                        double *xia, double *xka, double *jia, double *jka,
                        double *kia, double *kka, double *jij, double *jkj,
                        double *kij, double *kkj,
                        double *dintc1, double *dintx1, double *t1v1,
                        double *dintc2, double *dintx2, double *t1v2);

/* Same contract under the Fortran body's symbol: the argument list of
 * ccsd_trpdrv_omp_fbody, omp/ccsd_trpdrv_omp_f77.f:1-9 (every argument by
 * reference), bound by the USE_FORTRAN harness at omp/test.c:24-37. */
void ccsd_trpdrv_omp_fbody_(double *f1n, double *f1t, double *f2n, double *f2t,
                            double *f3n, double *f3t, double *f4n, double *f4t,
                            double *eorb, int *ncor, int *nocc, int *nvir,
                            double *emp4, double *emp5,
                            int *a, int *i, int *j, int *k, int *klo,
                            double *tij, double *tkj, double *tia, double *tka,
                            double *xia, double *xka, double *jia, double *jka,
                            double *kia, double *kka, double *jij, double *jkj,
                            double *kij, double *kkj,
                            double *dintc1, double *dintx1, double *t1v1,
                            double *dintc2, double *dintx2, double *t1v2);

/* Energy reduction over nvir*nvir (b,c) pairs of given f matrices; the four
 * outputs are overwritten, not accumulated.  eaijk, ncor, nocc, nvir by value.
 * Replaces ccsd_tengy_omp, omp/ccsd_tengy_c99.c:5-14. */
void ccsd_tengy_omp(const double *f1n, const double *f1t,
                    const double *f2n, const double *f2t,
                    const double *f3n, const double *f3t,
                    const double *f4n, const double *f4t,
                    const double *dintc1, const double *dintx1, const double *t1v1,
                    const double *dintc2, const double *dintx2, const double *t1v2,
                    const double *eorb, const double eaijk,
                    double *emp4i, double *emp5i, double *emp4k, double *emp5k,
                    const int ncor, const int nocc, const int nvir);

/* Same contract under the OpenCL variant's name, ocl/ccsd_tengy.c:3-12. */
void ccsd_tengy(const double *f1n, const double *f1t,
                const double *f2n, const double *f2t,
                const double *f3n, const double *f3t,
                const double *f4n, const double *f4t,
                const double *dintc1, const double *dintx1, const double *t1v1,
                const double *dintc2, const double *dintx2, const double *t1v2,
                const double *eorb, const double eaijk,
                double *emp4i, double *emp5i, double *emp4k, double *emp5k,
                const int ncor, const int nocc, const int nvir);

/* Fortran argument order and by-reference scalars: dintc1,dintc2,dintx1,
 * dintx2,t1v1,t1v2 and eaijk,ncor,nocc,nvir by reference.  Replaces the
 * Fortran ccsd_tengy_omp (symbol ccsd_tengy_omp_), omp/ccsd_tengy_f77.f:1-5. */
void ccsd_tengy_omp_(const double *f1n, const double *f1t,
                     const double *f2n, const double *f2t,
                     const double *f3n, const double *f3t,
                     const double *f4n, const double *f4t,
                     const double *dintc1, const double *dintc2,
                     const double *dintx1, const double *dintx2,
                     const double *t1v1, const double *t1v2,
                     const double *eorb, const double *eaijk,
                     double *emp4i, double *emp5i, double *emp4k, double *emp5k,
                     const int *ncor, const int *nocc, const int *nvir);

/* ccsd_tengy_omp's energy reduction on device-resident arrays, enqueued on
 * `stream` (a hipStream_t; null = the default stream), allocating nothing: for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * serving loops that forbid hipMalloc.  No counterpart in the reference.
 * f1n..f4t (nvir x nvir each, column major), the six vectors (nvir each) and
 * eorb (ncor+nocc+nvir) are device pointers with ccsd_tengy_omp's meaning; out4
 * (device, 4 doubles) receives emp4i, emp5i, emp4k, emp5k when the stream gets
 * there, bitwise what ccsd_tengy_omp returns for the same values.  The per-tile
 * partials go to a static device array of 66,560 bytes per GPU, held until the
 * process ends, so nvir is at most 2048 (CCSDT_E_ARG above).  Every call on one
 * GPU shares that array: calls that overlap on one GPU (other streams, other
 * threads) race.  Returns CCSDT_OK, CCSDT_E_ARG or CCSDT_E_DEVICE. */
int ccsdt_tengy_device_noalloc(const double *f1n, const double *f1t,
                               const double *f2n, const double *f2t,
                               const double *f3n, const double *f3t,
                               const double *f4n, const double *f4t,
                               const double *dintc1, const double *dintx1, const double *t1v1,
                               const double *dintc2, const double *dintx2, const double *t1v2,
                               const double *eorb, double eaijk, double *out4,
                               int ncor, int nocc, int nvir, void *stream);

/* ---------------------------------------------------------------- status */
/* No counterpart in the reference, which has no error reporting
 * (omp/ccsd_trpdrv_omp_cbody.c returns void and validates nothing). */
enum {
    CCSDT_OK = 0,
    CCSDT_E_ARG = -1,       /* invalid argument (sizes, indices, null pointer) */
    CCSDT_E_DEVICE = -2,    /* no usable gfx950 device or a HIP call failed */
    CCSDT_E_NOMEM = -3,     /* device or pinned allocation failed */
    CCSDT_E_BUDGET = -4,    /* serving admission: footprint above the budget */
    CCSDT_E_STATE = -5,     /* call out of sequence (destroyed job, ...) */
    CCSDT_E_INJECTED = -6   /* failure forced by ccsdt_debug_inject_failure */
};
int ccsdt_last_error(void);                /* last status of this thread */
const char *ccsdt_last_error_string(void); /* message for it */
void ccsdt_clear_error(void);
/* policy 1 (default): drop-ins print the message and abort() on failure.
 * policy 0: drop-ins record the error, set emp4 and emp5 (or the four tengy
 * outputs) to NaN and return. */
void ccsdt_set_error_policy(int abort_on_error);
/* Host-side fault injection for failure-path tests: the n-th (1-based) device
 * allocation or launch from now fails with CCSDT_E_INJECTED; 0 disables. */
void ccsdt_debug_inject_failure(int nth);
const char *ccsdt_version(void);
/* Selects and initialises a device (default 0).  Returns CCSDT_OK or an error. */
int ccsdt_init(int device);
/* When 0, the drop-in body skips the copy of f1n..f4t back to the host (the
 * reference leaves them filled; default 1).  Also set by CCSDT_F_WRITEBACK=0. */
void ccsdt_set_f_writeback(int on);

/* ---------------------------------------------------------------- inputs */
/* Seeded synthetic inputs: PCG32 (XSH-RR, 64-bit state, 32-bit output; O'Neill
 * 2014) on stream `stream`; doubles from 53 bits of two draws, scaled into
 * [lo,hi).  No counterpart: replaces make_array, omp/test.c:55-62. */
void ccsdt_pcg_raw(uint64_t seed, uint64_t stream, uint32_t *out, int64_t n);
void ccsdt_pcg_uniform(uint64_t seed, uint64_t stream, double lo, double hi,
                       double *out, int64_t n);

/* ---------------------------------------------------------------- job API */
/* Order of the arrays in ccsdt_job_desc.arrays: the body's argument order. */
enum {
    CCSDT_EORB = 0, CCSDT_TIJ, CCSDT_TKJ, CCSDT_TIA, CCSDT_TKA, CCSDT_XIA,
    CCSDT_XKA, CCSDT_JIA, CCSDT_JKA, CCSDT_KIA, CCSDT_KKA, CCSDT_JIJ,
    CCSDT_JKJ, CCSDT_KIJ, CCSDT_KKJ, CCSDT_DINTC1, CCSDT_DINTX1, CCSDT_T1V1,
    CCSDT_DINTC2, CCSDT_DINTX2, CCSDT_T1V2, CCSDT_NARRAYS
};

/* A job holds the harness-shaped arrays of omp/test.c:112-146: eorb (nbf =
 * ncor+nocc+nvir); tij, jia, kia (nvir^2); tkj, jka, kka (kchunk*nvir^2, slice
 * k-klo); tia, xia, jij, kij (nocc*nvir); tka, xka, jkj, kkj (kchunk*nocc*nvir,
 * slice k-klo); dintc1..t1v2 (nvir).  The arrays are copied to the device at
 * creation; host buffers may be reused afterwards. */
typedef struct {
    int ncor, nocc, nvir, kchunk;
    const double *arrays[CCSDT_NARRAYS];
} ccsdt_job_desc;

typedef struct ccsdt_job_s *ccsdt_job;

/* Arithmetic of the K=nvir contractions. */
enum {
    CCSDT_MODE_FP64 = 0,        /* fused FP64 MFMA GEMM + energy epilogue */
    CCSDT_MODE_FP64_SPLIT = 1,  /* FP64 GEMM writes f to HBM, separate tengy pass */
    CCSDT_MODE_OZAKI_INT8 = 2,  /* Ozaki INT8 slices, INT32 accumulate, FP64 recombine */
    CCSDT_MODE_BF16X3 = 3,      /* three BF16 planes, FP32 accumulate per level, FP64 recombine */
    CCSDT_MODE_FP32 = 4         /* FP32 throughout, as the FP32 body gpu-omp/ccsd_trpdrv_omp_cbody.c
                                   (operands rounded to FP32, FP32 MFMA GEMMs, FP32 energy
                                   summand); per-pair partials summed in FP64.  A precision
                                   comparison point */
};

/* Run flags. */
enum {
    CCSDT_RUN_PERSISTENT = 1,  /* fused kernel on a persistent grid (every workgroup slot
                                  of the device, each looping over work items) */
    CCSDT_RUN_GRAPH = 2        /* capture the run (uploads, launches, downloads) in a
                                  hipGraph once per signature and replay it */
};

/* Tile variants of the fused kernels.  FP64 modes: ONE_WAVE or TWO_WAVE.  The
 * split and FP32 modes run their two-wave block-pair kernel for 0 or TWO_WAVE;
 * Ozaki INT8 with at most 4 slices also runs ROW_PAIR. */
enum {
    CCSDT_TILE_ONE_WAVE = 0,  /* one wave computes all 16 products of a 16x16 block pair */
    CCSDT_TILE_TWO_WAVE = 1,  /* two waves share the pair, eight products each */
    CCSDT_TILE_ROW_PAIR = 2,  /* four waves share two block pairs stacked in one column
                                 block (16x16 blocks (r, c) and (r+1, c)), one product
                                 group of four each: fewer operand loads per matrix
                                 operation.  The f matrices are bitwise those of
                                 TWO_WAVE; the energy epilogue sums in another order, so
                                 terms agree to about 1e-13 of a triple's largest term */
    CCSDT_TILE_AUTO = -1      /* servers only: the tuning cache's pick (ccsdt_server_opts) */
};

/* Order in which the fused kernel's work items (triple x 16x16 block pair) run.
 * Results do not depend on it (each partial is stored by triple and pair). */
enum {
    CCSDT_ORDER_TASK = 0,  /* all block pairs of a triple together, triples in list order */
    CCSDT_ORDER_PAIR = 1   /* one block pair across all triples (sorted by k slice) at a
                              time: the workgroups in flight share operand panels */
};

typedef struct {
    int mode;       /* CCSDT_MODE_* */
    int slices;     /* Ozaki slice count (3..7), ignored otherwise; 0 = 4 */
    int flags;      /* CCSDT_RUN_* */
    int variant;    /* CCSDT_TILE_* */
    int wg_per_cu;  /* persistent grid: workgroups per CU (0 = as many as fit) */
    int order;      /* CCSDT_ORDER_* (fused modes; the separate-pass mode runs task order) */
    int occ_digits; /* CCSDT_MODE_OZAKI_INT8 with slices 3: 1 runs the K=nocc half as INT8
                       digit planes too, with row scales of its own (faster; adds that
                       half's split error), 0 as an FP64 tail.  Larger slice counts reject
                       1 (their kernels have no registers left for it); other modes
                       ignore it */
    int reserved;
} ccsdt_run_opts;

typedef struct {
    double emp4, emp5;          /* compensated sums over the list, i!=k rule applied */
    double ms_total;            /* device time of the whole run (HIP events) */
    double ms_pack;             /* operand packing */
    double ms_main;             /* GEMM (+ fused energy) kernel */
    double ms_tengy;            /* separate energy pass (MODE_FP64_SPLIT) */
    double ms_reduce;           /* per-triple and per-list reductions */
    int64_t ntriples;
    int64_t launches;           /* kernel launches of the run */
} ccsdt_run_result;

int ccsdt_job_create(const ccsdt_job_desc *desc, ccsdt_job *out);
/* Runs n triples (5 ints each: a,i,j,k,klo, 1-based as in the body).  terms
 * (may be NULL) receives 4*n doubles: emp4i, emp5i, emp4k, emp5k per triple.
 * fout (may be NULL, only for n == 1) receives the eight f matrices in the
 * body's order f1n,f1t,f2n,f2t,f3n,f3t,f4n,f4t. */
int ccsdt_job_run(ccsdt_job job, const int *triples, int64_t n,
                  const ccsdt_run_opts *opts, ccsdt_run_result *res,
                  double *terms, double *fout);
int ccsdt_job_destroy(ccsdt_job job);

/* Autotuner.  No counterpart in the reference (one DGEMM call per product,
 * omp/ccsd_trpdrv_omp_cbody.c:71-116, with MKL's own blocking).  Times every
 * launch configuration of the job's triple list -- item order x tile variant x
 * schedule (one workgroup per item, or persistent grids at the occupancy limit
 * and at half of it) --
 * with base's mode, slices and graph flag, and writes the fastest to *best
 * (candidates within 1% of the fastest count as equal; the first of them in the
 * order order x tile x schedule is taken, so repeated tunings agree; pair-major
 * order comes first).
 * Each candidate runs once untimed, then `reps` times; its time is the median
 * of ms_total.  table (may be NULL) receives up to cap entries, *count the
 * number of candidates tried.  Candidates the device rejects are recorded with
 * their status and skipped.  With mode CCSDT_MODE_OZAKI_INT8 and slices 0 the
 * precision is chosen first: the first of (S = 3 with occ_digits), S = 3, 4, .., 7
 * (cheapest first) whose E[T] = emp4 + emp5 over these triples is within 1e-9
 * relative of the FP64 form's (ten times under the split form's 1e-8 bound);
 * best->slices and best->occ_digits report it. */
typedef struct {
    ccsdt_run_opts opts;
    double ms;      /* median ms_total, or -1 if the candidate failed */
    int status;
    int pad;
} ccsdt_tune_entry;

int ccsdt_job_autotune(ccsdt_job job, const int *triples, int64_t n, const ccsdt_run_opts *base,
                       int reps, ccsdt_run_opts *best, ccsdt_tune_entry *table, int cap,
                       int *count);
/* Tuning cache.  ccsdt_job_autotune keeps every pick, keyed by the request (base's
 * mode, slices -- 0 meaning "choose" -- occ_digits and graph flag), the job (kind,
 * ncor, nocc, nvir, kchunk), the triple list (length and a hash of its quintuples),
 * the device (architecture, CU count) and the library build (its compile time:
 * another build's picks are not reused).  For a key it has seen it returns the
 * kept pick at once, without timing: *count = 0 and table untouched.  Picks are
 * kept in memory and, when a file is set, appended to it as lines of text.
 * ccsdt_tune_cache_file sets that file (NULL or "" for memory only) and reads
 * the picks it holds, returning how many; the environment variable
 * CCSDT_TUNE_CACHE sets it at the first lookup unless this was called.
 * ccsdt_tune_cache_clear forgets the in-memory picks (the file is kept);
 * ccsdt_tune_cache_size counts them.  No counterpart in the reference, which
 * leaves blocking to its BLAS (omp/ccsd_trpdrv_omp_cbody.c:71-116). */
int ccsdt_tune_cache_file(const char *path);
void ccsdt_tune_cache_clear(void);
int ccsdt_tune_cache_size(void);
/* Device bytes the job holds (arrays + packed operands + workspace for n triples). */
int64_t ccsdt_job_footprint(int ncor, int nocc, int nvir, int kchunk, int64_t ntriples);

/* ---------------------------------------------------------------- full-tensor jobs
 * No counterpart in the reference: its harness holds one slice per operand
 * (omp/test.c:112-146) and its NWChem caller fetches the slices of every
 * (i,j,k) from global arrays before each body call.  A full-tensor job holds the
 * slices of every occupied index, so any triple runs without a host gather.
 * With 0-based i, j, k, vv = nvir*nvir and ov = nocc*nvir, the body's operands
 * of triple (a,i,j,k) are (each slice in the body's layout):
 *   Tij = t2 + (j*nocc + i)*vv      Tkj = t2 + (j*nocc + k)*vv
 *   Jia = jv + i*vv                 Jka = jv + k*vv        (Kia, Kka from kv alike)
 *   Tia = t1o + i*ov                Tka = t1o + k*ov       (Xia, Xka from xo alike)
 *   Jij = jo + (j*nocc + i)*ov      Jkj = jo + (j*nocc + k)*ov  (Kij, Kkj from ko alike)
 * eorb and the six nvir vectors are as in ccsdt_job_desc (the same six for
 * every triple, as the harness passes them, omp/test.c:141-146,167-176).
 * Triples are run with ccsdt_job_run; their klo is checked (1 <= klo <= k)
 * and otherwise unused.
 *
 * Driver tensors (optional: t1, dja, djx all set, or all null).  The reference
 * allocates the driver's Dja, Djka, Djia (omp/test.c:137-139) but never uses
 * them; the six vectors of each triple are sliced from them by NWChem's
 * driver, which is not in the reference.  Given the driver tensors, the six
 * vector pointers above are not read and each triple reads, with rows of nvir
 * doubles and 0-based i, j, k:
 *   t1v1   = t1  + i*nvir                 t1v2   = t1  + k*nvir
 *   dintc1 = dja + (j*nocc + k)*nvir      dintc2 = dja + (j*nocc + i)*nvir
 *   dintx1 = djx + (j*nocc + k)*nvir      dintx2 = djx + (j*nocc + i)*nvir
 * so that for one j, dja + j*nocc*nvir has the harness's Dja shape (nvir x
 * nocc), djx + j*nocc*nvir its Djia, and that block's columns klo..khi its Djka
 * (nvir x kchunk).  This pairing (emp5i: t1 of i with integrals over j, k;
 * emp5k: t1 of k with integrals over j, i) keeps the i <-> k exchange symmetry
 * of omp/ccsd_tengy_c99.c:106-111; no reference code fixes it.  Nothing is
 * copied per triple: the kernels read the rows in place. */
typedef struct {
    int ncor, nocc, nvir, pad;
    const double *eorb;
    const double *t2, *jv, *kv, *t1o, *xo, *jo, *ko;
    const double *dintc1, *dintx1, *t1v1, *dintc2, *dintx2, *t1v2;
    const double *t1;   /* nocc*nvir: [m][c] */
    const double *dja;  /* nocc*nocc*nvir: [j][m][c] */
    const double *djx;  /* nocc*nocc*nvir: [j][m][c] */
} ccsdt_tensor_desc;

int ccsdt_job_create_tensors(const ccsdt_tensor_desc *desc, ccsdt_job *out);
/* driver = 1 for a job given the driver tensors. */
int64_t ccsdt_tensor_job_footprint(int ncor, int nocc, int nvir, int64_t ntriples, int driver);

/* ---------------------------------------------------------------- serving */
/* No counterpart in the reference (whose harness runs one job, omp/test.c).  A
 * server keeps admitted jobs device-resident and interleaves their triples in
 * shared launches ("ticks"): every tick takes a round-robin share of the
 * remaining triples of every admitted job, so jobs that arrive while others run
 * join the next tick.  Jobs are admitted in arrival order while the sum of their
 * device footprints stays under the budget; the others wait.  Each job's totals
 * are bitwise the same alone and in any mix (fixed per-job reduction order). */
typedef struct ccsdt_server_s *ccsdt_server;

/* ccsdt_server_opts.slices: choose the split form per job (Ozaki INT8). */
#define CCSDT_SLICES_PER_JOB (-1)

typedef struct {
    int64_t budget_bytes;    /* HBM budget for admitted jobs (0 = 200 GB) */
    int items_per_tick;      /* work items per tick (triple x 16x16 block pair); 0 = the
                                tuning cache's server pick for this mode and slice policy
                                on this device (ccsdt_server_tune_store), else 131072 */
    int mode;                /* CCSDT_MODE_FP64, CCSDT_MODE_OZAKI_INT8 or CCSDT_MODE_BF16X3 */
    int slices;              /* Ozaki slices (0 = 4), or CCSDT_SLICES_PER_JOB: each job's own
                                form, chosen at submit from its size (three slices with
                                occupied digit planes for a job of at least 1e8 energy
                                summands, triples x nvir^2, else four; the result reports
                                it); each tick then runs one launch per form */
    int variant;             /* CCSDT_TILE_* (FP64: ONE_WAVE or TWO_WAVE; Ozaki INT8 with
                                at most 4 slices may take ROW_PAIR), or CCSDT_TILE_AUTO: the
                                tuning cache's server pick, else 0 (as the job API's
                                default) */
    int occ_digits;          /* CCSDT_MODE_OZAKI_INT8 with slices 3: 1 packs every job's
                                K=nocc half as INT8 digit planes too (as
                                ccsdt_run_opts.occ_digits); 0 otherwise */
} ccsdt_server_opts;

typedef struct {
    double emp4, emp5;       /* totals with the i != k rule, compensated */
    double t_submit, t_admit, t_finish;  /* seconds on the server's clock */
    int64_t ntriples;
    int64_t footprint;       /* device bytes the job held */
    int status;              /* CCSDT_OK or the error that failed the job */
    int slices;              /* the job's Ozaki slice count (0 outside the INT8 form) */
    int occ_digits;          /* 1 when its K=nocc half ran as INT8 digit planes */
} ccsdt_job_result;

int ccsdt_server_create(const ccsdt_server_opts *opts, ccsdt_server *out);
/* Submits a job (arrays as for ccsdt_job_create, copied before return) and its
 * triples; *id receives the job id.  CCSDT_E_BUDGET if the job alone exceeds
 * the budget. */
int ccsdt_server_submit(ccsdt_server s, const ccsdt_job_desc *desc, const int *triples,
                        int64_t n, int64_t *id);
/* The same for a full-tensor job (footprint from ccsdt_tensor_job_footprint). */
int ccsdt_server_submit_tensors(ccsdt_server s, const ccsdt_tensor_desc *desc, const int *triples,
                                int64_t n, int64_t *id);
/* Retires finished ticks, finishes jobs, admits waiting jobs and launches ticks.
 * block = 1 waits for the oldest tick in flight first.  Returns the number of
 * jobs finished by this call, or a negative error. */
int ccsdt_server_progress(ccsdt_server s, int block);
/* 1 and *out filled when job id has finished, 0 while it is pending, negative
 * for an unknown id. */
int ccsdt_server_result(ccsdt_server s, int64_t id, ccsdt_job_result *out);
/* Jobs waiting, admitted (running), finished; device bytes in use. */
int ccsdt_server_counts(ccsdt_server s, int64_t *waiting, int64_t *running, int64_t *finished,
                        int64_t *bytes_in_use);
double ccsdt_server_clock(ccsdt_server s);  /* seconds since creation */
/* The options the server runs with: the caller's, with every default and cached
 * pick filled in. */
int ccsdt_server_options(ccsdt_server s, ccsdt_server_opts *out);
/* Records opts' tile (variant) and items_per_tick as the server pick for opts'
 * mode, slice policy and occupied digit planes on the current device, in the
 * tuning cache (memory, and the cache file when one is set); score_ms is kept
 * with it (the tuner's measure, e.g. p99 job latency).  Servers created later
 * with CCSDT_TILE_AUTO or items_per_tick = 0 use it.  No counterpart in the
 * reference: the serving loop has none. */
int ccsdt_server_tune_store(const ccsdt_server_opts *opts, double score_ms);
int ccsdt_server_destroy(ccsdt_server s);

/* ---------------------------------------------------------------- diagnostics */
/* One v_mfma_f64_16x16x4_f64 on A (16x4 row major, 64 doubles) and B (4x16 row
 * major, 64 doubles); d256[4*lane + r] receives each lane's four results
 * (D[lane/16 + 4r][lane%16] on gfx950).  No
 * counterpart in the reference: checks the lane layout the kernels assume. */
int ccsdt_selftest_mfma(const double *a64, const double *b64, double *d256);
/* One v_mfma_i32_16x16x64_i8 (shape 0; d receives 4 ints per lane),
 * v_mfma_i32_32x32x32_i8 (shape 1; 16 ints per lane) or
 * v_mfma_f32_16x16x32_bf16 (shape 2; the bits of 4 floats per lane) or
 * v_mfma_f32_16x16x4_f32 (shape 3; the first word of each lane's operand is the
 * float, 4 float results per lane) on raw operand registers: a256 and b256 hold
 * each lane's 16 operand bytes as four int32 words.  No counterpart in the
 * reference: checks the INT8, BF16 and FP32 lane layouts. */
int ccsdt_selftest_mfma_i8(const int32_t *a256, const int32_t *b256, int32_t *d, int shape);
/* Streams `bytes` through a plain device copy kernel `reps` times and returns the
 * average milliseconds per copy (or -1): the reference point for HBM counters.
 * No counterpart in the reference. */
double ccsdt_debug_copy(int64_t bytes, int reps);
/* Streams `bytes` through a plain device read-and-sum kernel (one warm-up launch,
 * then `reps` timed) and returns the average milliseconds per pass (or -1): the
 * HBM reference point for kernels that only read.  No counterpart in the
 * reference. */
double ccsdt_debug_read(int64_t bytes, int reps);
/* INT8 matrix-core rate on register operands, in TOP/s (or -1): shape 0 runs
 * v_mfma_i32_16x16x64_i8 (the split kernels' instruction), 1
 * v_mfma_i32_32x32x32_i8, at four waves per SIMD; random_operands = 1 feeds
 * per-lane pseudo-random bytes (consecutive MFMAs see different bits, as digit
 * planes do), 0 all ones.  The sustained rate depends on the operand bits.  No
 * counterpart in the reference. */
double ccsdt_debug_mfma_rate(int shape, int random_operands);
/* Nanoseconds per wave-wide load per CU when every CU streams L1-resident data
 * with loads of `bytes` (4, 8 or 16) per lane into VGPRs, or (to_lds = 1, 16
 * bytes) as global->LDS DMA; multiply by the shader clock for cycles.  No
 * counterpart in the reference: measures the per-instruction cost of the vector
 * memory path that bounds the split kernel's digit loop.  -1 on failure. */
double ccsdt_debug_load_rate(int bytes, int to_lds);
/* on = 0: jobs whose runs start after this call pack the K = nocc digit planes of
 * CCSDT_MODE_OZAKI_INT8 with occ_digits one digit per 64-wide K step even when
 * nocc <= 32 (the default, on = 1, puts two digits in each step there: fewer loads
 * and MFMAs, the same integer products, bitwise the same results).  For tests
 * that compare the two layouts.  No counterpart in the reference. */
void ccsdt_debug_occ_fold(int on);
/* Position of element (b, c) of f matrix m (0..7, body order f1n, f1t, ..,
 * f4t) inside one triple's 8 nv^2 doubles: in the device layout of the separate
 * energy pass (tiled = 1: 32x32 tile pairs one after the other, each pair's 16
 * tiles, the (b,c) tile and its (c,b) mirror of all eight matrices, in one
 * contiguous run) or in the reference's column-major layout (tiled = 0:
 * m nv^2 + b + c nv).  -1 for indices out of range.  Needs no device.  No
 * counterpart in the reference: for tests of the layout. */
int64_t ccsdt_debug_f_index(int m, int b, int c, int nv, int tiled);
/* Host-to-host copies of n buffers (dst[q] <- src[q], bytes[q] bytes) through the
 * copy pool the synchronous drop-ins stage their arrays and f matrices with (the
 * calling thread plus CCSDT_COPY_THREADS workers, default 4).  Needs no device.
 * Returns the pool's worker count, or -1 for a negative size or null pointer.
 * No counterpart in the reference (whose body reads host memory in place,
 * omp/ccsd_trpdrv_omp_cbody.c:71-116). */
int ccsdt_debug_host_copy(void *const *dst, const void *const *src, const int64_t *bytes, int n);

#ifdef __cplusplus
}
#endif
#endif /* CCSDT_H */
