// ccsdt_internal.h -- shared declarations between the HIP kernels and the
// host C-ABI layer of libccsdt.so.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccsdt {

// One (a,i,j,k) triple on the device: the four operand panel pairs it reads
// and its energy denominator shift.  Panel pair q is panels 2q and 2q+1 of the
// job's packed buffer:
//   li  = L(i)   = {LJ, LK}     (Jia/Tia, Kia/Xia)
//   lk  = L(k)   = {LJt, LKt}   (Jka/Tka, Kka/Xka)
//   rij = R(i,j) = {R1t, R3t}   (Tij with Kij, Jij)
//   rkj = R(k,j) = {R1, R3}     (Tkj with Kkj, Jkj)
// A harness-shaped job (omp/test.c arrays) has li = 0, rij = 1 and, for k
// slice s = k-klo (omp/ccsd_trpdrv_omp_cbody.c:64-70), rkj = 2+2s, lk = 3+2s;
// a full-tensor job indexes its L and R families by occupied indices.
// vrow[q] is the row (nvp doubles) of the job's vector pool holding vector q of
// the triple, in the kernels' order t1v1 t1v2 dintx1 dintx2 dintc1 dintc2: rows
// 0..5 when the job has one set of vectors, per-triple rows into the driver
// tensors otherwise (ccsdt_tensor_desc).
struct Task {
    int li, lk, rij, rkj;
    int ik_differ;  // 1 when i != k: the k terms enter the totals
    float eaijk32;  // the same in FP32 arithmetic on FP32 eorb (the FP32 form)
    double eaijk;   // eorb[a] - (eorb[ncor+i] + eorb[ncor+j] + eorb[ncor+k])
    int vrow[6];
    int pad;
};

enum { V_T1V1 = 0, V_T1V2 = 1, V_DINTX1 = 2, V_DINTX2 = 3, V_DINTC1 = 4, V_DINTC2 = 5 };

// Shapes of one job after padding.
struct Geom {
    int ncor, nocc, nvir;
    int nvp;     // nvir rounded up to 16 (MFMA block rows)
    int nb;      // nvp / 16
    int kdim;    // nvir + nocc: the concatenated contraction length
    int kp;      // kdim rounded up to 4
    int nks;     // kp / 4 MFMA contraction steps
    int npairs;  // nb*(nb+1)/2 lower-triangle 16x16 block pairs
    int nrpairs; // items per task of the row-pair tile: sum over ci of ceil((nb - ci) / 2)
    int64_t vecsz;  // doubles in one packed panel: nvp * kp
};

// Packed-panel descriptor: panel(x, kk) for x < nvir, kk < nvir + nocc is
//   kk <  nvir : first_t ? P[kk + x*nvir]  : P[x + kk*nvir]
//   kk >= nvir : sgn * (second_t ? Q[l + x*nocc] : Q[x + l*nvir]),  l = kk - nvir
// and 0 in the padding.
struct PanelDesc {
    const double *P;
    const double *Q;
    double *out;
    int first_t, second_t;
    double sgn;
};

// Panel order inside a harness-shaped job's packed buffer (pairs 0 and 1, then
// per k slice s the pairs 2+2s = {R1, R3} and 3+2s = {LJt, LKt}).
enum { PAN_LJ = 0, PAN_LK = 1, PAN_R1T = 2, PAN_R3T = 3, PAN_SLICE0 = 4 };
enum { SL_R1 = 0, SL_R3 = 1, SL_LJT = 2, SL_LKT = 3 };

// Geometry of the Ozaki INT8 split form: the K = nvir half of every panel as
// `slices` int8 digit planes (kv = nvir rounded up to 64, one
// v_mfma_i32_16x16x64_i8 K step per 64), and the K = nocc half either as an
// FP64 tail panel (kt = nocc rounded up to 8) or, at 3 slices with occupied
// digit planes, as more K steps of the same planes under row scales of their
// own (folded two digits per step when nocc <= 32).  BF16x3 uses the same
// fields with 32-wide K steps.
struct SplitGeom {
    int slices;
    int bf16;             // 1: three BF16 planes (v_mfma_f32_16x16x32_bf16), not INT8 digits
    int kv, nkv;          // kv = round_up(nvir, 64), nkv = kv / 64 (BF16: 32)
    int kx, nkx;          // INT8 with occupied digit planes (S <= kTail8MaxSlices): the K =
                          // nocc half as planes of its own, kx = round_up(nocc, 64),
                          // nkx = kx / 64; folded (below): kx = 128, nkx = 2; else 0
    int fold;             // 1: occupied digit planes with nocc <= 32, two digits per
                          // 64-wide K step (the folded slots of k_split_pack)
    int kt, nkt;          // FP64 K = nocc tail (when nkx = 0): kt = round_up(nocc, 8), nkt = kt / 4 (even)
    int64_t slice_bytes;  // bytes of one digit plane of one panel: nvp * (kv + kx) (BF16: 2 nvp kv)
    int64_t tail_sz;      // doubles of one FP64 tail panel: nvp * kt (0 when nkx > 0)
    int npanels;          // panels of the job (bounds of the buffer resources)
};

// INT8 slice counts whose K = nocc half may run as digit planes (own row scales,
// ccsdt_run_opts::occ_digits) rather than as an FP64 tail.  At S = 3 the digit
// step costs less than the FP64 MFMA tail; at S >= 4 its accumulators and
// operands would cost the kernel a wave per SIMD (more than 168 VGPRs).
constexpr int kTail8MaxSlices = 3;
// Added to an INT8 slice count in the launchers' `slices` code: the K = nocc half
// as digit planes (S <= kTail8MaxSlices only).
constexpr int kSlicesOccDigits = 16;

// Row-scale sets of a split job: the K = nvir planes' scales, then (nkx > 0) the
// K = nocc planes' scales, each npanels * nvp doubles.
inline int split_scale_sets(const SplitGeom &sg) { return sg.nkx > 0 ? 2 : 1; }

// Everything a kernel needs to know about one job (device-resident table).
struct JobDev {
    Geom g;
    SplitGeom sg;
    const double *pan;      // FP64 packed panels
    const int8_t *digits;   // split form: digit planes
    const double *scales;   // split form: per-row power-of-two scales
    const double *tail;     // split form: FP64 K = nocc panels
    const double *ev;       // virtual orbital energies, padded to nvp
    const double *vec6;     // vector pool: rows of nvp doubles (Task::vrow)
    const short2 *pairs;    // lower-triangle 16x16 block pairs
    const int4 *rpairs;     // row-pair tile items: (r0, ci, pair index of (r0, ci), pair
                            // index of (r0 + 1, ci) or -1 when r0 + 1 = nb)
};

// A run of consecutive tasks of one job inside a launch.  Work items are
// (task, block pair); the segment's items start at item0.  Task-major order
// (pmajor = 0): item L of the segment is pair L % npairs of task L / npairs.
// Pair-major order (pmajor = 1): item L is pair L / ntasks of task
// order[L % ntasks], so the workgroups in flight share one block pair (and,
// with order sorting the tasks by k slice, the same operand panels).  Partials
// are stored by (task, pair) either way, so the order changes no result.
struct Segment {
    const Task *tasks;   // the segment's first task
    double *partial;     // 4 doubles per (task, pair), task-major, from the first task
    double *fout;        // f matrices of the first task (fstride apart) or null
    double *terms;       // 4 doubles per task (emp4i, emp5i, emp4k, emp5k)
    const int *order;    // pair-major: execution order of the tasks (null = as stored)
    int job;             // index into the JobDev table
    int item0;           // first work item of the segment
    int ntasks;
    int pmajor;
    int ftiled;          // fout in the tiled layout of f_index (the separate energy
                         // pass); 0: column-major, as the reference's f1n..f4t
    int pad;
};

// Element (b, c) of f matrix m (body order f1n, f1t, ..., f4t) inside one
// triple's 8 nv^2 doubles.  Column-major, m nv^2 + b + c nv, is the reference's
// layout (f1n..f4t).  The tiled layout serves the separate energy pass, whose
// work unit is a 32x32 tile pair: the tile D at (block bi, block ci), bi >= ci,
// and its mirror M at (ci, bi), of all eight matrices.  It stores the pairs one
// after the other (bi = 0, 1, ..., ci = 0..bi), each as D of matrices 0..7 then
// M of matrices 0..7 (a diagonal pair has D only), each tile column-major; the
// last block row and column are narrower (w = nv - 32 (nb - 1)).  So a tile pair
// is one contiguous run of at most 128 KiB, and the 8 nv^2 doubles hold every
// element once: the pairs before row bi take 8 x 1024 bi^2 doubles, those
// before (bi, ci) in its row 8 x 64 w(bi) ci more.
__host__ __device__ inline int64_t f_index(int m, int b, int c, int nv, bool tiled)
{
    if (!tiled) return (int64_t)m * nv * nv + b + (int64_t)c * nv;
    const int tb = b >> 5, tc = c >> 5;
    const int bi = tb > tc ? tb : tc, ci = tb > tc ? tc : tb;
    const int wbi = nv - 32 * bi < 32 ? nv - 32 * bi : 32;
    const int wci = nv - 32 * ci < 32 ? nv - 32 * ci : 32;
    const int64_t base = 8 * ((int64_t)1024 * bi * bi + (int64_t)64 * wbi * ci);
    const int64_t sz = (int64_t)wbi * wci;
    if (tb >= tc)  // D (rows in block bi = tb, columns in block ci = tc)
        return base + m * sz + (b - 32 * tb) + (int64_t)(c - 32 * tc) * wbi;
    return base + (8 + m) * sz + (b - 32 * tb) + (int64_t)(c - 32 * tc) * wci;  // M
}

// Size of the current work list, read by the kernels (so a persistent launch
// of fixed shape can be replayed on new lists).
struct WorkHdr {
    int total;    // work items
    int nsegs;
    int ntasks;   // tasks over all segments
    int pad;
    // Persistent launches: items handed out so far from each of the 8 item
    // ranges (one per XCD).  The host writes zeros with every upload of the
    // header, so each launch (and each graph replay) starts from zero.
    int next[8];
};

// Launchers (ccsdt_kernels.hip).  All enqueue on `st` and return hipError_t.
// pairs: the step-pair panel layout the two-wave FP64 kernel reads (k_fused2);
// otherwise the single-step layout of k_fused.
hipError_t launch_pack(const PanelDesc *d_desc, int npanels, const Geom &g, bool pairs, hipStream_t st);
// grid = number of workgroups: the item count for a plain launch, or a
// multiple of 8 for a persistent one (each workgroup then loops over items).
// variant 0: one wave computes all 16 products of a block pair (k_fused);
// variant 1: two waves share the pair, eight products each (k_fused2).
hipError_t launch_fused(const JobDev *jobs, const Segment *segs, WorkHdr *hdr, int grid,
                        bool persist, int variant, int64_t fstride, bool tengy, bool writef,
                        hipStream_t st);
hipError_t launch_tengy(const double *f, int64_t fstride, bool tiled, const Geom &g, const Task *tasks,
                        int ntasks, const double *ev, const double *vec6, double *partial,
                        hipStream_t st);
// One triple's energy-pass inputs in the caller's device arrays: f1n, f1t, ...,
// f4t (nvir x nvir column major each); the virtual orbital energies, then t1v1,
// t1v2, dintx1, dintx2, dintc1, dintc2 (nvir each, unpadded); eaijk.
struct TengyDirect {
    const double *f[8];
    const double *v[7];
    double eaijk;
};
// Allocation-free energy pass (ccsdt_tengy_device_noalloc): its per-tile
// partials go to a static device array of kNoAllocPartials doubles, enough for
// nvir up to kNoAllocMaxNvir (64 x 65 / 2 tile pairs of 32 x 32, four doubles each).
constexpr int kNoAllocMaxNvir = 2048;
constexpr int kNoAllocPartials = (kNoAllocMaxNvir / 32) * (kNoAllocMaxNvir / 32 + 1) / 2 * 4;
// The current device's address of that array (hipGetSymbolAddress).
hipError_t scratch_address(double **ptr);
// The pass on `st` into the array at scratch_ptr, then the per-tile sums into
// out4 (device, 4 doubles: emp4i, emp5i, emp4k, emp5k).
hipError_t launch_tengy_direct(const TengyDirect &in, const Geom &g, double *scratch_ptr, double *out4,
                               hipStream_t st);
int tengy_npairs(const Geom &g);
hipError_t launch_split_pack(const PanelDesc *d_desc, int npanels, const Geom &g,
                             const SplitGeom &sg, int8_t *digits, double *scales, double *tail,
                             hipStream_t st);
// variant kTileRowPair: the row-pair tile (k_split_rows; INT8 with slices 3, 4 or
// 3 + kSlicesOccDigits); otherwise the two-wave block-pair kernels.
hipError_t launch_fused_split(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                              int grid, bool persist, int slices, int variant, int64_t fstride,
                              bool tengy, bool writef, hipStream_t st);
// Per-task sums of the pair partials over all segments (grid may be persistent).
hipError_t launch_reduce_segments(const JobDev *jobs, const Segment *segs, const WorkHdr *hdr,
                                  int grid, hipStream_t st);
// `slices` codes of launch_fused_split / fused_blocks_per_cu for the BF16x3
// form and the FP32 form (both two-wave block-pair kernels).
constexpr int kSlicesBf16x3 = 100;
constexpr int kSlicesFp32 = 200;
// FP32 panels (same element order as k_pack's, FP32 values; panel p at
// pan32 + p * vecsz).
hipError_t launch_pack32(const PanelDesc *d_desc, int npanels, const Geom &g, float *pan32,
                         hipStream_t st);
// Tile variant of the split kernels: two row blocks over one column block on four
// waves (CCSDT_TILE_ROW_PAIR in ccsdt.h).
constexpr int kTileRowPair = 2;
// Largest INT8 slice count the row-pair tile runs (its accumulators for two
// blocks must fit two waves per SIMD).
constexpr int kRowPairMaxSlices = 4;
// Largest K for which adjacent digit levels combine exactly in INT32,
// (acc_L << 7) + acc_(L+1) (S <= 4).  Per K term a level sum is at most 16129,
// 16256, 20352 and 24448 in magnitude (levels 0..3: |d_1| <= 127, later digits
// <= 64), so the largest combined value, levels 2 and 3 at S = 4, is at most
// (20352 x 128 + 24448) K = 2629504 K < 2^31 for K <= 816.
constexpr int kIntPairMaxK = 768;
// Largest K (nvir, and nocc) the row-pair tile runs: it always combines levels
// in INT32.
constexpr int kRowPairMaxK = kIntPairMaxK;
// Resident workgroups per CU of the fused kernels (for persistent grids).
int fused_blocks_per_cu(bool split, int slices, int variant);
hipError_t launch_reduce_terms(const double *partial, int npairs, int ntasks, double *terms,
                               hipStream_t st);
hipError_t launch_reduce_list(const double *terms, const Task *tasks, int ntasks, double *out2,
                              hipStream_t st);
hipError_t launch_mfma_probe(const double *a, const double *b, double *d, hipStream_t st);
hipError_t launch_mfma_i8_probe(const int *a, const int *b, int *d, int shape, hipStream_t st);
// INT8 MFMA rate probe: grid workgroups of 256 threads into out[grid * 256].
hipError_t launch_mfma_rate(int *out, int shape, bool rnd, int grid, int iters, hipStream_t st);
hipError_t launch_copy(const double *src, double *dst, int64_t n, hipStream_t st);
// Streaming read-and-sum of n doubles into out[blocks] (blocks <= kReadSumBlocks).
constexpr int kReadSumBlocks = 16384;
hipError_t launch_readsum(const double *src, double *out, int64_t n, int blocks, hipStream_t st);
// Load-rate probe (ccsdt_debug_load_rate): grid workgroups of 256 threads, each
// thread issuing iters*8 loads of `bytes` (4, 8, 16) per lane, to LDS when lds.
hipError_t launch_load_rate(const uint32_t *buf, uint32_t *out, int bytes, bool lds, int grid,
                            int iters, hipStream_t st);

}  // namespace ccsdt
