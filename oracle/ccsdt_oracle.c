/*
 * ccsdt_oracle.c -- CPU restatement of the CCSD(T) triples body and its energy
 * reduction.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and
 * the timed CPU baseline.  Nothing in the product library links, loads or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load liboracle.so.
 *
 * What it restates (reference files, read-only):
 *   orc_body       omp/ccsd_trpdrv_omp_cbody.c:31-145  (16 DGEMMs, eaijk, tengy,
 *                  i!=k accumulation rule)
 *   orc_tengy      omp/ccsd_tengy_c99.c:5-119 (32x32 tile loop, collapse(2)
 *                  OpenMP reduction, the per-(b,c) formula at :65-111)
 *   orc_run_list   omp/test.c:160-179 (sequential calls, per-call wall timer),
 *                  minus the +-1000 wrap of omp/test.c:188-191 (a harness hack,
 *                  not part of the body)
 *   orc_pcg_*      the seeded synthetic inputs (PCG32 XSH-RR 64/32, O'Neill 2014)
 *                  that replace omp/test.c:55-62 make_array.
 *
 * DGEMM is the reference's external BLAS (omp/ccsd_trpdrv_omp_cbody.c:3-18).  The
 * provider is picked at run time: MKL's cblas_dgemm through dlopen when the image
 * has libmkl_rt, else a plain column-major triple loop written here.  Both follow
 * the textbook definition C <- alpha*op(A)*op(B) + beta*C.
 *
 * Pinning (tests/test_oracle.py): tengy and the body glue are compared bit for
 * bit with the real reference sources compiled into oracle/_ref (single thread);
 * the GEMM operand mapping is checked against an exact rational brute force and
 * numpy on small shapes; PCG32 against its published known-answer outputs.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ PCG32 */
typedef struct { uint64_t state, inc; } orc_pcg_t;

static uint32_t orc_pcg_next(orc_pcg_t *g)
{
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

static void orc_pcg_seed(orc_pcg_t *g, uint64_t seed, uint64_t stream)
{
    g->state = 0u;
    g->inc = (stream << 1u) | 1u;
    orc_pcg_next(g);
    g->state += seed;
    orc_pcg_next(g);
}

/* n raw 32-bit outputs (known-answer test hook). */
void orc_pcg_raw(uint64_t seed, uint64_t stream, uint32_t *out, long n)
{
    orc_pcg_t g;
    orc_pcg_seed(&g, seed, stream);
    for (long i = 0; i < n; i++) out[i] = orc_pcg_next(&g);
}

/* Uniform doubles in [lo,hi): 53 random bits from two draws. */
static double orc_pcg_unit(orc_pcg_t *g)
{
    uint32_t a = orc_pcg_next(g) >> 5, b = orc_pcg_next(g) >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

void orc_pcg_uniform(uint64_t seed, uint64_t stream, double lo, double hi,
                     double *out, long n)
{
    orc_pcg_t g;
    orc_pcg_seed(&g, seed, stream);
    for (long i = 0; i < n; i++) out[i] = lo + (hi - lo) * orc_pcg_unit(&g);
}

/* ------------------------------------------------------------------ DGEMM */
enum { ORC_COLMAJOR = 102, ORC_NOTRANS = 111, ORC_TRANS = 112 };
typedef void (*orc_cblas_dgemm_t)(int, int, int, int, int, int, double,
                                  const double *, int, const double *, int,
                                  double, double *, int);
static orc_cblas_dgemm_t orc_mkl_dgemm = NULL;
static int orc_provider = 0; /* 0 = not chosen, 1 = own loop, 2 = MKL */

/* C <- alpha*op(A)*op(B) + beta*C, column major, op(A) m x k, op(B) k x n. */
static void orc_dgemm_loop(int ta, int tb, int m, int n, int k, double alpha,
                           const double *A, int lda, const double *B, int ldb,
                           double beta, double *C, int ldc)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        double *cj = C + (long)j * ldc;
        if (beta == 0.0) {
            for (int i = 0; i < m; i++) cj[i] = 0.0;
        } else if (beta != 1.0) {
            for (int i = 0; i < m; i++) cj[i] *= beta;
        }
        if (!ta) {
            for (int l = 0; l < k; l++) {
                double blj = tb ? B[j + (long)l * ldb] : B[l + (long)j * ldb];
                double s = alpha * blj;
                const double *al = A + (long)l * lda;
                for (int i = 0; i < m; i++) cj[i] += al[i] * s;
            }
        } else {
            for (int i = 0; i < m; i++) {
                double acc = 0.0;
                for (int l = 0; l < k; l++) {
                    double blj = tb ? B[j + (long)l * ldb] : B[l + (long)j * ldb];
                    acc += A[l + (long)i * lda] * blj;
                }
                cj[i] += alpha * acc;
            }
        }
    }
}

/* which: 0 auto (MKL if loadable), 1 own loop, 2 MKL only.  Returns provider. */
int orc_blas_select(int which)
{
    if (which == 1) { orc_provider = 1; return 1; }
    if (!orc_mkl_dgemm) {
        const char *env = getenv("CCSDT_ORACLE_MKL");
        const char *cands[] = { env, "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.1" };
        setenv("MKL_THREADING_LAYER", "GNU", 0);
        setenv("MKL_INTERFACE_LAYER", "LP64", 0);
        for (unsigned c = 0; c < sizeof(cands) / sizeof(cands[0]) && !orc_mkl_dgemm; c++) {
            if (!cands[c] || !cands[c][0]) continue;
            void *h = dlopen(cands[c], RTLD_NOW | RTLD_LOCAL);
            if (h) orc_mkl_dgemm = (orc_cblas_dgemm_t)dlsym(h, "cblas_dgemm");
        }
    }
    if (orc_mkl_dgemm) { orc_provider = 2; return 2; }
    if (which == 2) return -1;
    orc_provider = 1;
    return 1;
}

const char *orc_blas_name(void)
{
    return orc_provider == 2 ? "mkl" : orc_provider == 1 ? "oracle-loop" : "unset";
}

void orc_dgemm(int ta, int tb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc)
{
    if (!orc_provider) orc_blas_select(0);
    if (orc_provider == 2)
        orc_mkl_dgemm(ORC_COLMAJOR, ta ? ORC_TRANS : ORC_NOTRANS,
                      tb ? ORC_TRANS : ORC_NOTRANS, m, n, k, alpha, A, lda, B, ldb,
                      beta, C, ldc);
    else
        orc_dgemm_loop(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* ------------------------------------------------------------------ tengy */
/* Per-(b,c) summand of the energy reduction, omp/ccsd_tengy_c99.c:65-111.
 * bc = b + c*nvir and cb = c + b*nvir address the same column-major f matrices;
 * the expression keeps the reference's association so a single-threaded run is
 * bitwise comparable. */
#define ORC_TILE 32

void orc_tengy(const double *f1n, const double *f1t, const double *f2n, const double *f2t,
               const double *f3n, const double *f3t, const double *f4n, const double *f4t,
               const double *dintc1, const double *dintx1, const double *t1v1,
               const double *dintc2, const double *dintx2, const double *t1v2,
               const double *eorb, double eaijk,
               double *emp4i_, double *emp5i_, double *emp4k_, double *emp5k_,
               int ncor, int nocc, int nvir, int nthreads)
{
    double s4i = 0.0, s5i = 0.0, s4k = 0.0, s5k = 0.0;
    const double *ev = eorb + ncor + nocc; /* virtual orbital energies */
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for collapse(2) num_threads(nthreads) reduction(+ : s4i, s5i, s4k, s5k)
    for (int b0 = 0; b0 < nvir; b0 += ORC_TILE) {
        for (int c0 = 0; c0 < nvir; c0 += ORC_TILE) {
            int bend = b0 + ORC_TILE < nvir ? b0 + ORC_TILE : nvir;
            int cend = c0 + ORC_TILE < nvir ? c0 + ORC_TILE : nvir;
            for (int b = b0; b < bend; b++) {
                for (int c = c0; c < cend; c++) {
                    const double d = -1.0 / (ev[b] + ev[c] + eaijk);
                    const long bc = b + (long)c * nvir, cb = c + (long)b * nvir;
                    const double n1 = f1n[bc], t1 = f1t[bc], n1r = f1n[cb], t1r = f1t[cb];
                    const double n2 = f2n[bc], t2 = f2t[bc], n2r = f2n[cb], t2r = f2t[cb];
                    const double n3 = f3n[bc], t3 = f3t[bc], n3r = f3n[cb], t3r = f3t[cb];
                    const double n4 = f4n[bc], t4 = f4t[bc], n4r = f4n[cb], t4r = f4t[cb];
                    s4i += d * (t1 + n1r + t2r + n3 + n4r) * (t1 - t2 * 2 - t3 * 2 + t4)
                         - d * (n1 + t1r + n2r + n3r) * (t1 * 2 - t2 - t3 + t4 * 2)
                         + d * 3 * (n1 * (n1 + n3r + t4r * 2) + n2 * t2r + n3 * t4);
                    s4k += d * (n1 + t1r + n2r + t3 + t4r) * (n1 - n2 * 2 - n3 * 2 + n4)
                         - d * (t1 + n1r + t2r + t3r) * (n1 * 2 - n2 - n3 + n4 * 2)
                         + d * 3 * (t1 * (t1 + t3r + n4r * 2) + t2 * n2r + t3 * n4);
                    const double u1 = t1v1[b], u2 = t1v2[b];
                    const double x1 = dintx1[c], x2 = dintx2[c];
                    const double y1 = dintc1[c], y2 = dintc2[c];
                    s5i += d * u1 * x1 * (t1 + n2 + n4r - (t3 + n4 + n2r + n1 + t2 + n3r) * 2
                                          + (n3 + t4 + n1r) * 4)
                         + d * u1 * y1 * (n1 + n4 + t1r - (n2 + n3 + t2r) * 2);
                    s5k += d * u2 * x2 * (n1 + t2 + t4r - (n3 + t4 + t2r + t1 + n2 + t3r) * 2
                                          + (t3 + n4 + t1r) * 4)
                         + d * u2 * y2 * (t1 + t4 + n1r - (t2 + t3 + n2r) * 2);
                }
            }
        }
    }
    *emp4i_ = s4i;
    *emp5i_ = s5i;
    *emp4k_ = s4k;
    *emp5k_ = s5k;
}

/* Sum of absolute summands of each term: the scale used for relative errors of
 * terms that cancel (emp5 in particular). */
void orc_tengy_abs(const double *f1n, const double *f1t, const double *f2n, const double *f2t,
                   const double *f3n, const double *f3t, const double *f4n, const double *f4t,
                   const double *dintc1, const double *dintx1, const double *t1v1,
                   const double *dintc2, const double *dintx2, const double *t1v2,
                   const double *eorb, double eaijk, double *abs4 /* [4] */,
                   int ncor, int nocc, int nvir)
{
    double a4i = 0, a5i = 0, a4k = 0, a5k = 0;
    const double *ev = eorb + ncor + nocc;
#pragma omp parallel for reduction(+ : a4i, a5i, a4k, a5k)
    for (int b = 0; b < nvir; b++) {
        for (int c = 0; c < nvir; c++) {
            const double d = -1.0 / (ev[b] + ev[c] + eaijk);
            const long bc = b + (long)c * nvir, cb = c + (long)b * nvir;
            const double n1 = f1n[bc], t1 = f1t[bc], n1r = f1n[cb], t1r = f1t[cb];
            const double n2 = f2n[bc], t2 = f2t[bc], n2r = f2n[cb], t2r = f2t[cb];
            const double n3 = f3n[bc], t3 = f3t[bc], n3r = f3n[cb], t3r = f3t[cb];
            const double n4 = f4n[bc], t4 = f4t[bc], n4r = f4n[cb], t4r = f4t[cb];
            a4i += fabs(d * (t1 + n1r + t2r + n3 + n4r) * (t1 - t2 * 2 - t3 * 2 + t4))
                 + fabs(d * (n1 + t1r + n2r + n3r) * (t1 * 2 - t2 - t3 + t4 * 2))
                 + fabs(d * 3 * (n1 * (n1 + n3r + t4r * 2) + n2 * t2r + n3 * t4));
            a4k += fabs(d * (n1 + t1r + n2r + t3 + t4r) * (n1 - n2 * 2 - n3 * 2 + n4))
                 + fabs(d * (t1 + n1r + t2r + t3r) * (n1 * 2 - n2 - n3 + n4 * 2))
                 + fabs(d * 3 * (t1 * (t1 + t3r + n4r * 2) + t2 * n2r + t3 * n4));
            a5i += fabs(d * t1v1[b] * dintx1[c] * (t1 + n2 + n4r - (t3 + n4 + n2r + n1 + t2 + n3r) * 2 + (n3 + t4 + n1r) * 4))
                 + fabs(d * t1v1[b] * dintc1[c] * (n1 + n4 + t1r - (n2 + n3 + t2r) * 2));
            a5k += fabs(d * t1v2[b] * dintx2[c] * (n1 + t2 + t4r - (n3 + t4 + t2r + t1 + n2 + t3r) * 2 + (t3 + n4 + t1r) * 4))
                 + fabs(d * t1v2[b] * dintc2[c] * (t1 + t4 + n1r - (t2 + t3 + n2r) * 2));
        }
    }
    abs4[0] = a4i; abs4[1] = a5i; abs4[2] = a4k; abs4[3] = a5k;
}

/* ------------------------------------------------------------------ body */
/* The eight f matrices of one (a,i,j,k) triple, omp/ccsd_trpdrv_omp_cbody.c:71-116.
 * Each f is a K=nvir product (beta=0) followed by a K=nocc update (alpha=-1,
 * beta=1).  k and klo are 1-based; slices of the k-indexed operands start at
 * (k-klo)*nvir^2 or (k-klo)*nocc*nvir. */
void orc_fmats(double *f1n, double *f1t, double *f2n, double *f2t,
               double *f3n, double *f3t, double *f4n, double *f4t,
               int nocc, int nvir, int k1, int klo1,
               const double *tij, const double *tkj, const double *tia, const double *tka,
               const double *xia, const double *xka, const double *jia, const double *jka,
               const double *kia, const double *kka, const double *jij, const double *jkj,
               const double *kij, const double *kkj)
{
    const int nv = nvir, no = nocc;
    const long vv = (long)nvir * nvir, ov = (long)nocc * nvir;
    const long s = (long)(k1 - klo1);
    const double *tkj_k = tkj + s * vv, *jka_k = jka + s * vv, *kka_k = kka + s * vv;
    const double *kkj_k = kkj + s * ov, *jkj_k = jkj + s * ov;
    const double *tka_k = tka + s * ov, *xka_k = xka + s * ov;
    /* "n" side: (a,i) operands against the (k,j) slice */
    orc_dgemm(0, 1, nv, nv, nv, 1.0, jia, nv, tkj_k, nv, 0.0, f1n, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, tia, nv, kkj_k, no, 1.0, f1n, nv);
    orc_dgemm(0, 1, nv, nv, nv, 1.0, kia, nv, tkj_k, nv, 0.0, f2n, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, xia, nv, kkj_k, no, 1.0, f2n, nv);
    orc_dgemm(0, 0, nv, nv, nv, 1.0, jia, nv, tkj_k, nv, 0.0, f3n, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, tia, nv, jkj_k, no, 1.0, f3n, nv);
    orc_dgemm(0, 0, nv, nv, nv, 1.0, kia, nv, tkj_k, nv, 0.0, f4n, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, xia, nv, jkj_k, no, 1.0, f4n, nv);
    /* "t" side: (k,a) slice against the (i,j) operands */
    orc_dgemm(0, 1, nv, nv, nv, 1.0, jka_k, nv, tij, nv, 0.0, f1t, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, tka_k, nv, kij, no, 1.0, f1t, nv);
    orc_dgemm(0, 1, nv, nv, nv, 1.0, kka_k, nv, tij, nv, 0.0, f2t, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, xka_k, nv, kij, no, 1.0, f2t, nv);
    orc_dgemm(0, 0, nv, nv, nv, 1.0, jka_k, nv, tij, nv, 0.0, f3t, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, tka_k, nv, jij, no, 1.0, f3t, nv);
    orc_dgemm(0, 0, nv, nv, nv, 1.0, kka_k, nv, tij, nv, 0.0, f4t, nv);
    orc_dgemm(0, 0, nv, nv, no, -1.0, xka_k, nv, jij, no, 1.0, f4t, nv);
}

/* eaijk, omp/ccsd_trpdrv_omp_cbody.c:122-126: a is an absolute 1-based orbital
 * index, i,j,k are 1-based occupied indices offset by ncor. */
double orc_eaijk(const double *eorb, int ncor, int a1, int i1, int j1, int k1)
{
    return eorb[a1 - 1] - (eorb[ncor + i1 - 1] + eorb[ncor + j1 - 1] + eorb[ncor + k1 - 1]);
}

/* One body call with the reference's argument list; terms4 (may be NULL)
 * receives emp4i, emp5i, emp4k, emp5k. */
void orc_body(double *f1n, double *f1t, double *f2n, double *f2t,
              double *f3n, double *f3t, double *f4n, double *f4t,
              double *eorb, int *ncor_, int *nocc_, int *nvir_,
              double *emp4_, double *emp5_,
              int *a_, int *i_, int *j_, int *k_, int *klo_,
              double *tij, double *tkj, double *tia, double *tka,
              double *xia, double *xka, double *jia, double *jka,
              double *kia, double *kka, double *jij, double *jkj,
              double *kij, double *kkj,
              double *dintc1, double *dintx1, double *t1v1,
              double *dintc2, double *dintx2, double *t1v2,
              double *terms4, int nthreads)
{
    const int ncor = *ncor_, nocc = *nocc_, nvir = *nvir_;
    double e4i, e5i, e4k, e5k;
    orc_fmats(f1n, f1t, f2n, f2t, f3n, f3t, f4n, f4t, nocc, nvir, *k_, *klo_,
              tij, tkj, tia, tka, xia, xka, jia, jka, kia, kka, jij, jkj, kij, kkj);
    const double eaijk = orc_eaijk(eorb, ncor, *a_, *i_, *j_, *k_);
    orc_tengy(f1n, f1t, f2n, f2t, f3n, f3t, f4n, f4t, dintc1, dintx1, t1v1,
              dintc2, dintx2, t1v2, eorb, eaijk, &e4i, &e5i, &e4k, &e5k,
              ncor, nocc, nvir, nthreads);
    double emp4 = *emp4_, emp5 = *emp5_;
    emp4 += e4i;
    emp5 += e5i;
    if (*i_ != *k_) { /* omp/ccsd_trpdrv_omp_cbody.c:136-139 */
        emp4 += e4k;
        emp5 += e5k;
    }
    *emp4_ = emp4;
    *emp5_ = emp5;
    if (terms4) { terms4[0] = e4i; terms4[1] = e5i; terms4[2] = e4k; terms4[3] = e5k; }
}

/* ------------------------------------------------------------------ list driver */
/* Array order of the job tables (matches the body's argument order). */
enum { J_EORB, J_TIJ, J_TKJ, J_TIA, J_TKA, J_XIA, J_XKA, J_JIA, J_JKA, J_KIA, J_KKA,
       J_JIJ, J_JKJ, J_KIJ, J_KKJ, J_DINTC1, J_DINTX1, J_T1V1, J_DINTC2, J_DINTX2,
       J_T1V2, J_NARR };

/* Runs the triples (a,i,j,k,klo, 1-based, five ints each) in list order, as
 * omp/test.c:160-179 does.  Outputs (each may be NULL): terms[4*n] per-triple
 * emp4i,emp5i,emp4k,emp5k; abs4[4*n] absolute-summand scales; secs[n] per-call
 * wall time; tot[4] = { plain running emp4, emp5 (reference accumulation),
 * compensated emp4, emp5 }.  Stops early after max_seconds (>0) and returns
 * the number of triples done. */
long orc_run_list(int ncor, int nocc, int nvir, const double *const *arr,
                  const int *trip, long n, double *terms, double *abs4,
                  double *secs, double *tot, double max_seconds, int nthreads)
{
    const long vv = (long)nvir * nvir;
    double *f = (double *)malloc(sizeof(double) * 8 * vv);
    if (!f) return -1;
    double emp4 = 0.0, emp5 = 0.0;
    double c4 = 0.0, c4c = 0.0, c5 = 0.0, c5c = 0.0; /* Neumaier sums */
    const double t_start = omp_get_wtime();
    long done = 0;
    for (long t = 0; t < n; t++) {
        int nc = ncor, no = nocc, nv = nvir;
        int a = trip[5 * t], i = trip[5 * t + 1], j = trip[5 * t + 2];
        int k = trip[5 * t + 3], klo = trip[5 * t + 4];
        double tm[4];
        double t0 = omp_get_wtime();
        orc_body(f, f + vv, f + 2 * vv, f + 3 * vv, f + 4 * vv, f + 5 * vv, f + 6 * vv,
                 f + 7 * vv, (double *)arr[J_EORB], &nc, &no, &nv, &emp4, &emp5,
                 &a, &i, &j, &k, &klo,
                 (double *)arr[J_TIJ], (double *)arr[J_TKJ], (double *)arr[J_TIA],
                 (double *)arr[J_TKA], (double *)arr[J_XIA], (double *)arr[J_XKA],
                 (double *)arr[J_JIA], (double *)arr[J_JKA], (double *)arr[J_KIA],
                 (double *)arr[J_KKA], (double *)arr[J_JIJ], (double *)arr[J_JKJ],
                 (double *)arr[J_KIJ], (double *)arr[J_KKJ],
                 (double *)arr[J_DINTC1], (double *)arr[J_DINTX1], (double *)arr[J_T1V1],
                 (double *)arr[J_DINTC2], (double *)arr[J_DINTX2], (double *)arr[J_T1V2],
                 tm, nthreads);
        double t1 = omp_get_wtime();
        if (secs) secs[t] = t1 - t0;
        if (terms) memcpy(terms + 4 * t, tm, sizeof tm);
        if (abs4) {
            const double eaijk = orc_eaijk(arr[J_EORB], ncor, a, i, j, k);
            orc_tengy_abs(f, f + vv, f + 2 * vv, f + 3 * vv, f + 4 * vv, f + 5 * vv,
                          f + 6 * vv, f + 7 * vv, arr[J_DINTC1], arr[J_DINTX1],
                          arr[J_T1V1], arr[J_DINTC2], arr[J_DINTX2], arr[J_T1V2],
                          arr[J_EORB], eaijk, abs4 + 4 * t, ncor, nocc, nvir);
        }
        double add4[2] = { tm[0], i != k ? tm[2] : 0.0 };
        double add5[2] = { tm[1], i != k ? tm[3] : 0.0 };
        for (int q = 0; q < 2; q++) {
            double s = c4 + add4[q];
            c4c += fabs(c4) >= fabs(add4[q]) ? (c4 - s) + add4[q] : (add4[q] - s) + c4;
            c4 = s;
            s = c5 + add5[q];
            c5c += fabs(c5) >= fabs(add5[q]) ? (c5 - s) + add5[q] : (add5[q] - s) + c5;
            c5 = s;
        }
        done = t + 1;
        if (max_seconds > 0 && t1 - t_start > max_seconds) break;
    }
    if (tot) { tot[0] = emp4; tot[1] = emp5; tot[2] = c4 + c4c; tot[3] = c5 + c5c; }
    free(f);
    return done;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
