// ccsdt_kernels.hip -- gfx950 kernels of the CCSD(T) triples body.
//
// The sixteen DGEMMs of one (a,i,j,k) triple (omp/ccsd_trpdrv_omp_cbody.c:71-116)
// are regrouped as dot products of length K = nvir + nocc between eight packed
// operand vectors per virtual index x:
//   LJ(x)  = [Jia(x,:)  Tia(x,:)]      R1(x)  = [Tkj(x,:)  -Kkj(:,x)]
//   LK(x)  = [Kia(x,:)  Xia(x,:)]      R3(x)  = [Tkj(:,x)  -Jkj(:,x)]
//   LJt(x) = [Jka(x,:)  Tka(x,:)]      R1t(x) = [Tij(x,:)  -Kij(:,x)]
//   LKt(x) = [Kka(x,:)  Xka(x,:)]      R3t(x) = [Tij(:,x)  -Jij(:,x)]
// so that f1n(b,c) = LJ(b).R1(c), f2n = LK.R1, f3n = LJ.R3, f4n = LK.R3 and
// f1t = LJt.R1t, f2t = LKt.R1t, f3t = LJt.R3t, f4t = LKt.R3t.  The energy
// reduction (omp/ccsd_tengy_c99.c:65-111) at (b,c) needs every f at (b,c) and
// at (c,b); a wave that owns the 16x16 block pair (B,C) therefore computes the
// eight products at (b,c) and the eight mirrored ones f(c,b) = L(c).R(b) with
// the operand roles swapped, which puts all sixteen values of a (b,c) point in
// the same lane and register, and finishes the energy terms in registers.
//
// Panels are stored pre-swizzled for v_mfma_f64_16x16x4_f64: per 16-row block
// xb and 4-wide contraction step ks, lane l holds
//   value(x = 16*xb + l%16, kk = 4*ks + l/16),
// the A (or B) operand of one MFMA, so a wave's operand load is one contiguous
// 512-byte global_load_dwordx2.
#include <type_traits>

#include "ccsdt_internal.h"

namespace ccsdt {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

template <typename T>
struct SumT {
    T e4i, e5i, e4k, e5k;
};
typedef SumT<double> Sum4;

// Energy summand of one (b,c) point.  n*, t* are f?n, f?t at (b,c); nr*, tr*
// at (c,b); d = -1/(e_b + e_c + eaijk); u = t1v(b); x = dintx(c); y = dintc(c).
// The association follows omp/ccsd_tengy_c99.c:91-111 term for term; products
// feeding an add are fused (FMA).  That changes each point's rounding by about
// one ulp of its summands (far below the 1e-11 parity bound) and cuts the
// epilogue's FP64 instructions by about a quarter: 1.5 % of the split kernel's
// time at nocc=30, nvir=300 (the FP64 kernel, bound by its MFMAs, is unchanged).
// T = float evaluates the same operations in FP32, as the FP32 body of
// gpu-omp/ccsd_trpdrv_omp_cbody.c:204-221 does.
template <typename T>
__device__ __forceinline__ void tengy_point(T n1, T n2, T n3, T n4, T t1, T t2, T t3, T t4,
                                            T nr1, T nr2, T nr3, T nr4, T tr1, T tr2, T tr3,
                                            T tr4, T d, T u1, T u2, T x1, T x2, T y1, T y2,
                                            SumT<T> &s)
{
#pragma clang fp contract(fast)
    s.e4i += d * (t1 + nr1 + tr2 + n3 + nr4) * (t1 - t2 * 2 - t3 * 2 + t4)
           - d * (n1 + tr1 + nr2 + nr3) * (t1 * 2 - t2 - t3 + t4 * 2)
           + d * 3 * (n1 * (n1 + nr3 + tr4 * 2) + n2 * tr2 + n3 * t4);
    s.e4k += d * (n1 + tr1 + nr2 + t3 + tr4) * (n1 - n2 * 2 - n3 * 2 + n4)
           - d * (t1 + nr1 + tr2 + tr3) * (n1 * 2 - n2 - n3 + n4 * 2)
           + d * 3 * (t1 * (t1 + tr3 + nr4 * 2) + t2 * nr2 + t3 * n4);
    s.e5i += d * u1 * x1 * (t1 + n2 + nr4 - (t3 + n4 + nr2 + n1 + t2 + nr3) * 2
                            + (n3 + t4 + nr1) * 4)
           + d * u1 * y1 * (n1 + n4 + tr1 - (n2 + n3 + tr2) * 2);
    s.e5k += d * u2 * x2 * (n1 + t2 + tr4 - (n3 + t4 + tr2 + t1 + n2 + tr3) * 2
                            + (t3 + n4 + tr1) * 4)
           + d * u2 * y2 * (t1 + t4 + nr1 - (t2 + t3 + nr2) * 2);
}

// A 32-bit lane move through DPP (data-parallel primitives: the operand is read
// from another lane of the same row of 16 as part of the VALU instruction).
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

typedef int int2v __attribute__((ext_vector_type(2)));

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const int2v h = __builtin_bit_cast(int2v, v);
    const int2v r = {dpp_i32<CTRL>(h.x), dpp_i32<CTRL>(h.y)};
    return __builtin_bit_cast(double, r);
}

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ double lane_read(double v, int l)
{
    const int2v h = __builtin_bit_cast(int2v, v);
    const int2v r = {__builtin_amdgcn_readlane(h.x, l), __builtin_amdgcn_readlane(h.y, l)};
    return __builtin_bit_cast(double, r);
}

__device__ __forceinline__ float lane_read(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Sum over the wave, returned in every lane; every lane must be active.  Within
// each row of 16 lanes DPP moves add the lanes of each quad (quad_perm [1,0,3,2]
// and [2,3,0,1]) and then the row's quads (row_ror 4 and 8); the four row sums
// are read from lanes 0, 16, 32 and 48.  No LDS traffic (a shuffle is a
// ds_bpermute per 32-bit half and step, each waited for), and a fixed order of
// the additions, so the sum is the same on every run.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x124>(v);
    v += dpp_mov<0x128>(v);
    return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}

// Raw buffer addressing: a 128-bit resource in scalar registers, a wave-uniform
// byte offset (soffset) and one per-lane byte offset, so an operand fetch costs
// no per-load vector address registers.  Word 3 = 0x00020000 is the gfx9-family
// raw-buffer format (32-bit data format, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const double *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)bytes,
                                             0x00020000);
}

__device__ __forceinline__ double buffer_load_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  uint32_t soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Loads through a global (not flat) pointer: a flat load would also count
// against the LDS counter, tying every LDS wait to its latency.
__device__ __forceinline__ dbl2 gload2(const double *p)
{
    return *(const __attribute__((address_space(1))) dbl2 *)p;
}

__device__ __forceinline__ double gld(const double *p)
{
    return *(const __attribute__((address_space(1))) double *)p;
}

// Two doubles per lane in one 16-byte load.  On gfx950 a wave's 8-byte load
// takes as long in the texture data path as a 16-byte one (about 16 cycles per
// CU), so the step-pair panel layout below halves that path's time per MFMA.
__device__ __forceinline__ dbl2 buffer_load_f64x2(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                  uint32_t soff)
{
    return __builtin_bit_cast(dbl2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// Spreads consecutive logical tiles over one XCD: the hardware hands block b to
// XCD b % 8, so logical tile L = (b%8)*q + min(b%8, r) + b/8 keeps a task's
// tiles (which share operand panels) in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int total)
{
    const int q = total >> 3, r = total & 7;
    const int x = bid & 7, idx = bid >> 3;
    return x * q + min(x, r) + idx;
}

// Work schedule shared by the plain and the persistent launches: the items are
// split into 8 contiguous ranges, one per XCD (the hardware deals workgroup b to
// XCD b % 8).  A plain launch (one workgroup per item) maps workgroup b to item
// start(b%8) + b/8, the same mapping as xcd_remap, so every XCD works through
// its range in dispatch order.  A persistent grid (a multiple of 8 workgroups)
// keeps that order with a counter per range: each workgroup takes the next
// item of range b%8 with an atomic add until the range is used up.  (A static
// stride let the workgroups drift apart over thousands of items, widening the
// set of tasks in flight past what one XCD's L2 holds.)
struct ItemRange {
    int start, len;  // the range of XCD b % 8
};

__device__ __forceinline__ ItemRange item_range(int total, int b)
{
    const int x = b & 7;
    const int q = total >> 3, r = total & 7;
    return ItemRange{x * q + min(x, r), q + (x < r ? 1 : 0)};
}

// Next item of this workgroup's range for a persistent launch, or -1 when the
// range is used up (wave-uniform; one atomic per workgroup, lane 0 of wave 0).
// NW = waves per workgroup; with two waves the value passes through LDS, and
// the caller's item work must end with a barrier before the next call.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// A read of memory no kernel writes during the launch (tasks, execution order,
// block pairs: all copied from the host before it) through the constant
// address space, so a wave-uniform address becomes a scalar load.
template <typename T>
__device__ __forceinline__ T cread(const T *p)
{
    return *(const __attribute__((address_space(4))) T *)p;
}

// A task record and a block pair read field by field with cread (an aggregate
// copy would drop the address space).
__device__ __forceinline__ Task read_task(const Task *p)
{
    Task t;
    t.li = cread(&p->li);
    t.lk = cread(&p->lk);
    t.rij = cread(&p->rij);
    t.rkj = cread(&p->rkj);
    t.ik_differ = cread(&p->ik_differ);
    t.eaijk32 = cread(&p->eaijk32);
    t.eaijk = cread(&p->eaijk);
#pragma unroll
    for (int q = 0; q < 6; ++q) t.vrow[q] = cread(&p->vrow[q]);
    t.pad = 0;
    return t;
}

__device__ __forceinline__ short2 read_pair(const short2 *p)
{
    const int v = cread(reinterpret_cast<const int *>(p));
    return make_short2((short)(v & 0xffff), (short)(v >> 16));
}

// Item of workgroup b in a plain launch (grid = item count), or -1.
__device__ __forceinline__ int plain_item(const ItemRange &ir)
{
    const int j = blockIdx.x >> 3;
    return uniform(j < ir.len ? ir.start + j : -1);
}

template <int NW>
__device__ __forceinline__ int next_item(WorkHdr *hdr, const ItemRange &ir)
{
    int c = 0;
    if (threadIdx.x == 0) c = atomicAdd(&hdr->next[blockIdx.x & 7], 1);
    if (NW == 1) {
        c = __builtin_amdgcn_readlane(c, 0);
    } else {
        __shared__ int s_next;
        if (threadIdx.x == 0) s_next = c;
        __syncthreads();
        c = __builtin_amdgcn_readfirstlane(s_next);
    }
    return c < ir.len ? ir.start + c : -1;
}


// Task and block pair of work item `item` of segment sg (wave-uniform).
__device__ __forceinline__ void decode_item(const Segment &sg, int npairs, int item, int &t, int &p)
{
    const int L = uniform(item - sg.item0);
    if (sg.pmajor) {
        p = uniform(L / sg.ntasks);
        const int tt = L - p * sg.ntasks;
        t = uniform(sg.order ? cread(sg.order + tt) : tt);
    } else {
        t = uniform(L / npairs);
        p = uniform(L - t * npairs);
    }
}

// Segment holding work item `item` (segments sorted by item0; wave-uniform).
__device__ __forceinline__ int find_segment(const Segment *__restrict__ segs, int nsegs, int item)
{
    int lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].item0 <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ packing
// Panel element order.  Single-step layout (k_fused): per 16-row block and
// 4-wide K step, element 64*step + l holds row l%16, K index 4*step + l/16.
// Step-pair layout (PAIRS, the two-wave kernels): the block's steps go in pairs
// (2s, 2s+1) and element 128*s + 2*l + h holds lane l's value of step 2s+h, so
// one 16-byte load per lane feeds two steps; an odd last step keeps the
// single-step order.  A block is nks*64 doubles either way.
template <bool PAIRS>
__global__ __launch_bounds__(256) void k_pack(const PanelDesc *__restrict__ descs, Geom g)
{
    const PanelDesc d = descs[blockIdx.y];
    const int nv = g.nvir, no = g.nocc;
    const int64_t blk = (int64_t)g.nks * 64;
    const int64_t paired = (int64_t)(g.nks & ~1) * 64;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < g.vecsz;
         idx += (int64_t)gridDim.x * 256) {
        int l, ks, xb;
        if (PAIRS) {
            xb = (int)(idx / blk);
            const int64_t e = idx - (int64_t)xb * blk;
            if (e < paired) {
                ks = 2 * (int)(e >> 7) + (int)(e & 1);
                l = (int)((e >> 1) & 63);
            } else {
                ks = (int)(e >> 6);
                l = (int)(e & 63);
            }
        } else {
            l = (int)(idx & 63);
            const int64_t rest = idx >> 6;
            ks = (int)(rest % g.nks);
            xb = (int)(rest / g.nks);
        }
        const int x = xb * 16 + (l & 15);
        const int kk = ks * 4 + (l >> 4);
        double v = 0.0;
        if (x < nv) {
            if (kk < nv) {
                v = d.first_t ? d.P[kk + (int64_t)x * nv] : d.P[x + (int64_t)kk * nv];
            } else if (kk < nv + no) {
                const int l2 = kk - nv;
                v = d.sgn * (d.second_t ? d.Q[l2 + (int64_t)x * no] : d.Q[x + (int64_t)l2 * nv]);
            }
        }
        d.out[idx] = v;
    }
}

hipError_t launch_pack(const PanelDesc *d_desc, int npanels, const Geom &g, bool pairs, hipStream_t st)
{
    int64_t blocks = (g.vecsz + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (pairs)
        hipLaunchKernelGGL(k_pack<true>, dim3((unsigned)blocks, npanels), dim3(256), 0, st, d_desc, g);
    else
        hipLaunchKernelGGL(k_pack<false>, dim3((unsigned)blocks, npanels), dim3(256), 0, st, d_desc, g);
    return hipGetLastError();
}

// FP32 panels for the FP32 form: k_pack's element order with the values
// rounded to FP32 (operand lane layout of v_mfma_f32_16x16x4_f32, which
// matches the FP64 one: lane l holds row l%16, K index l/16 of a 4-wide step).
__global__ __launch_bounds__(256) void k_pack32(const PanelDesc *__restrict__ descs, Geom g,
                                                float *__restrict__ pan32)
{
    const PanelDesc d = descs[blockIdx.y];
    float *out = pan32 + (int64_t)blockIdx.y * g.vecsz;
    const int nv = g.nvir, no = g.nocc;
    for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < g.vecsz;
         idx += (int64_t)gridDim.x * 256) {
        const int l = (int)(idx & 63);
        const int64_t rest = idx >> 6;
        const int ks = (int)(rest % g.nks);
        const int xb = (int)(rest / g.nks);
        const int x = xb * 16 + (l & 15);
        const int kk = ks * 4 + (l >> 4);
        float v = 0.f;
        if (x < nv) {
            if (kk < nv) {
                v = (float)(d.first_t ? d.P[kk + (int64_t)x * nv] : d.P[x + (int64_t)kk * nv]);
            } else if (kk < nv + no) {
                const int l2 = kk - nv;
                v = (float)d.sgn * (float)(d.second_t ? d.Q[l2 + (int64_t)x * no] : d.Q[x + (int64_t)l2 * nv]);
            }
        }
        out[idx] = v;
    }
}

hipError_t launch_pack32(const PanelDesc *d_desc, int npanels, const Geom &g, float *pan32,
                         hipStream_t st)
{
    int64_t blocks = (g.vecsz + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_pack32, dim3((unsigned)blocks, npanels), dim3(256), 0, st, d_desc, g, pan32);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fused GEMM + energy
#define MF(p, ia, ib) acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ia], b[ib], acc[p], 0, 0, 0)

// K loop of k_fused.  Ping-pong operand buffers X and Y, two K steps per trip:
// each buffer's loads are in flight while the other buffer's MFMAs run; an odd
// last step runs from X after the loop.  DIAG skips the eight mirrored products.
// Resources of a task's four panel pairs, indexed by these constants.
enum { RP_LI = 0, RP_LK = 1, RP_RIJ = 2, RP_RKJ = 3 };

// Panel pairs of one task, wave-uniform (Task fields after readfirstlane).
struct PanelPairs {
    int li, lk, rij, rkj;
};

__device__ __forceinline__ PanelPairs panel_pairs(const Task &tk)
{
    return PanelPairs{uniform(tk.li), uniform(tk.lk), uniform(tk.rij), uniform(tk.rkj)};
}

// One raw-buffer resource per panel pair (two panels each), so byte offsets fit
// 32 bits whatever the job's panel count.
__device__ __forceinline__ void pair_resources(__amdgpu_buffer_rsrc_t (&rp)[4], const double *pan,
                                               int64_t vecsz, const PanelPairs &pp)
{
    const uint32_t bytes = (uint32_t)(2 * vecsz * 8);
    rp[RP_LI] = buffer_rsrc(pan + 2 * (int64_t)pp.li * vecsz, bytes);
    rp[RP_LK] = buffer_rsrc(pan + 2 * (int64_t)pp.lk * vecsz, bytes);
    rp[RP_RIJ] = buffer_rsrc(pan + 2 * (int64_t)pp.rij * vecsz, bytes);
    rp[RP_RKJ] = buffer_rsrc(pan + 2 * (int64_t)pp.rkj * vecsz, bytes);
}

template <bool DIAG>
__device__ __forceinline__ void fused_mainloop(dbl4 (&acc)[16], const __amdgpu_buffer_rsrc_t (&rp)[4],
                                               const uint32_t (&AO)[8], const uint32_t (&BO)[8],
                                               uint32_t voff, int nks)
{
    // the panel pair each operand lives in; A side (index b): LJ, LK, LJt, LKt,
    // R1, R3, R1t, R3t; B side (index c): R1, R3, R1t, R3t, LJ, LK, LJt, LKt
    constexpr int AR[8] = {RP_LI, RP_LI, RP_LK, RP_LK, RP_RKJ, RP_RKJ, RP_RIJ, RP_RIJ};
    constexpr int BR[8] = {RP_RKJ, RP_RKJ, RP_RIJ, RP_RIJ, RP_LI, RP_LI, RP_LK, RP_LK};
    constexpr int NV = DIAG ? 4 : 8;
    double xa[8], xb[8], ya[8], yb[8];
#define STEP_LOAD(da, db, kstep)                                                              \
    _Pragma("unroll") for (int v = 0; v < NV; ++v) {                                        \
        da[v] = buffer_load_f64(rp[AR[v]], voff, AO[v] + (uint32_t)(kstep) * 512);           \
        db[v] = buffer_load_f64(rp[BR[v]], voff, BO[v] + (uint32_t)(kstep) * 512);           \
    }
#define STEP_MFMA(da, db)                                                                     \
    {                                                                                         \
        double *a = da, *b = db;                                                              \
        MF(0, 0, 0);  MF(1, 1, 0);  MF(2, 0, 1);  MF(3, 1, 1);                                \
        MF(4, 2, 2);  MF(5, 3, 2);  MF(6, 2, 3);  MF(7, 3, 3);                                \
        if (!DIAG) {                                                                          \
            MF(8, 4, 4);  MF(9, 4, 5);  MF(10, 5, 4); MF(11, 5, 5);                           \
            MF(12, 6, 6); MF(13, 6, 7); MF(14, 7, 6); MF(15, 7, 7);                           \
        }                                                                                     \
    }
    STEP_LOAD(xa, xb, 0)
    int ks = 0;
    for (; ks + 1 < nks; ks += 2) {
        STEP_LOAD(ya, yb, ks + 1)
        STEP_MFMA(xa, xb)
        const int k2 = (ks + 2 < nks) ? ks + 2 : nks - 1;  // clamped prefetch, in bounds
        STEP_LOAD(xa, xb, k2)
        STEP_MFMA(ya, yb)
    }
    if (nks & 1) STEP_MFMA(xa, xb)
#undef STEP_LOAD
#undef STEP_MFMA
}

// One work item of k_fused: the 16x16 block pair (bi, ci) of one task.
template <bool TENGY, bool WRITEF>
__device__ __forceinline__ void fused_item(const JobDev &jd, const Task &tk, const PanelPairs &pp,
                                           int bi, int ci, int lane, double *__restrict__ part4,
                                           double *__restrict__ fo, bool ftiled)
{
    const Geom &g = jd.g;
    const double *__restrict__ ev = jd.ev;
    const double *__restrict__ vec6 = jd.vec6;
    __amdgpu_buffer_rsrc_t rp[4];
    pair_resources(rp, jd.pan, g.vecsz, pp);
    const uint32_t vsb = (uint32_t)(g.vecsz * 8);
    const uint32_t blkb = (uint32_t)g.nks * 512;  // bytes of one 16-row block of a panel
    const uint32_t ob = bi * blkb, oc = ci * blkb;
    const uint32_t voff = lane * 8;
    // operand v is the first (v even) or second panel of its pair
    uint32_t AO[8], BO[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        AO[v] = (v & 1) * vsb + ob;
        BO[v] = (v & 1) * vsb + oc;
    }
    dbl4 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = dbl4{0.0, 0.0, 0.0, 0.0};

    // Diagonal pairs (B = C) compute only the eight products at (b,c): their
    // mirrored values are the same products at (c,b), transposed through LDS below.
    const bool diag = (bi == ci);
    if (diag)
        fused_mainloop<true>(acc, rp, AO, BO, voff, g.nks);
    else
        fused_mainloop<false>(acc, rp, AO, BO, voff, g.nks);
    if (diag) {
        // f(c,b) at (b,c) = product q at (c,b): element (row, col) sits in lane
        // col + 16*(row%4), register row/4 (result layout below)
        __shared__ double tr[16][17];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
#pragma unroll
            for (int r = 0; r < 4; ++r) tr[(lane >> 4) + 4 * r][lane & 15] = acc[q][r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[8 + q][r] = tr[lane & 15][(lane >> 4) + 4 * r];
            __syncthreads();
        }
    }

    // v_mfma_f64_16x16x4_f64 result: lane l, element r holds row l/16 + 4*r,
    // column l%16 (measured with ccsdt_selftest_mfma).
    const int nv = g.nvir, nvp = g.nvp;
    const int col = lane & 15, row0 = lane >> 4;
    int vo[6];  // offsets of the task's six vectors in the pool (uniform)
#pragma unroll
    for (int q = 0; q < 6; ++q) vo[q] = uniform(tk.vrow[q] * nvp);
    const int cx = ci * 16 + col;
    if (WRITEF) {
        const int mat[8] = {0, 2, 4, 6, 1, 3, 5, 7};  // product -> body order f1n,f1t,...
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int bx = bi * 16 + row0 + 4 * r;
            if (bx < nv && cx < nv) {
#pragma unroll
                for (int q = 0; q < 8; ++q) fo[f_index(mat[q], bx, cx, nv, ftiled)] = acc[q][r];
                if (bi != ci) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) fo[f_index(mat[q], cx, bx, nv, ftiled)] = acc[8 + q][r];
                }
            }
        }
    }
    if (TENGY) {
        Sum4 s = {0.0, 0.0, 0.0, 0.0};
        const double ec = gld(ev + cx);
        const double u1c = gld(vec6 + vo[0] + cx), u2c = gld(vec6 + vo[1] + cx);
        const double x1c = gld(vec6 + vo[2] + cx), x2c = gld(vec6 + vo[3] + cx);
        const double y1c = gld(vec6 + vo[4] + cx), y2c = gld(vec6 + vo[5] + cx);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int bx = bi * 16 + row0 + 4 * r;
            if (bx < nv && cx < nv) {
                const double d = -1.0 / (gld(ev + bx) + ec + tk.eaijk);
                const double u1b = gld(vec6 + vo[0] + bx), u2b = gld(vec6 + vo[1] + bx);
                tengy_point(acc[0][r], acc[1][r], acc[2][r], acc[3][r],
                            acc[4][r], acc[5][r], acc[6][r], acc[7][r],
                            acc[8][r], acc[9][r], acc[10][r], acc[11][r],
                            acc[12][r], acc[13][r], acc[14][r], acc[15][r],
                            d, u1b, u2b, x1c, x2c, y1c, y2c, s);
                if (bi != ci) {  // the (c,b) point of the same pair of blocks
                    const double x1b = gld(vec6 + vo[2] + bx), x2b = gld(vec6 + vo[3] + bx);
                    const double y1b = gld(vec6 + vo[4] + bx), y2b = gld(vec6 + vo[5] + bx);
                    tengy_point(acc[8][r], acc[9][r], acc[10][r], acc[11][r],
                                acc[12][r], acc[13][r], acc[14][r], acc[15][r],
                                acc[0][r], acc[1][r], acc[2][r], acc[3][r],
                                acc[4][r], acc[5][r], acc[6][r], acc[7][r],
                                d, u1c, u2c, x1b, x2b, y1b, y2b, s);
                }
            }
        }
        s.e4i = wave_sum(s.e4i);
        s.e5i = wave_sum(s.e5i);
        s.e4k = wave_sum(s.e4k);
        s.e5k = wave_sum(s.e5k);
        if (lane == 0) {
            part4[0] = s.e4i; part4[1] = s.e5i; part4[2] = s.e4k; part4[3] = s.e5k;
        }
    }
}
#undef MF

// PERSIST: the workgroup loops over its items (persistent grid); otherwise the
// grid has one workgroup per item and no loop (the loop costs registers).
template <bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(64) void k_fused(const JobDev *__restrict__ jobs,
                                              const Segment *__restrict__ segs,
                                              WorkHdr *hdr, int64_t fstride)
{
    const int nsegs = hdr->nsegs;
    const ItemRange ir = item_range(hdr->total, blockIdx.x);
    for (int item = PERSIST ? next_item<1>(hdr, ir) : plain_item(ir); item >= 0;
         item = PERSIST ? next_item<1>(hdr, ir) : -1) {
        const int si = uniform(find_segment(segs, nsegs, item));
        const Segment sg = segs[si];
        const JobDev jd = jobs[uniform(sg.job)];
        int t, p;
        decode_item(sg, jd.g.npairs, item, t, p);
        const Task tk = read_task(sg.tasks + t);
        const short2 pr = read_pair(jd.pairs + p);
        fused_item<TENGY, WRITEF>(jd, tk, panel_pairs(tk), uniform(pr.x), uniform(pr.y), threadIdx.x,
                                  sg.partial + ((int64_t)t * jd.g.npairs + p) * 4,
                                  WRITEF ? sg.fout + (int64_t)t * fstride : nullptr, sg.ftiled != 0);
    }
}

static hipError_t launch_fused2(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                                int grid, bool persist, int64_t fstride, bool tengy, bool writef,
                                hipStream_t st);

hipError_t launch_fused(const JobDev *jobs, const Segment *segs, WorkHdr *hdr, int grid,
                        bool persist, int variant, int64_t fstride, bool tengy, bool writef,
                        hipStream_t st)
{
    if (grid <= 0) return hipErrorInvalidValue;
    if (variant == 1) return launch_fused2(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    if (variant != 0) return hipErrorInvalidValue;
    dim3 g3((unsigned)grid), block(64);
    if (persist && tengy && !writef)
        hipLaunchKernelGGL((k_fused<true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (persist)
        return hipErrorInvalidValue;
    else if (tengy && writef)
        hipLaunchKernelGGL((k_fused<true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (tengy)
        hipLaunchKernelGGL((k_fused<true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (writef)
        hipLaunchKernelGGL((k_fused<false, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ------------------------------------------------------------------ Ozaki INT8 split form
// The K = nvir half of every product runs as Ozaki-I split products (Ozaki,
// Ogita, Oishi, Rump 2012; integer form after Ootomo, Ozaki, Yokota 2024): each
// panel row x is scaled by 2^-e(x) so its largest entry is at most 127/128, then
// cut into `S` signed 7-bit digit planes by round-to-nearest,
//   P(x,k) = 2^e(x) * sum_s d_s(x,k) 2^(-7s) + r,  |d_1| <= 127, |d_s>1| <= 64,
// every step exact in FP64.  Digit plane s of A times plane q of B is an exact
// INT32 product (K*127^2 << 2^31); products with s+q <= S+1 are kept and summed
// per level L = s+q in INT32, then recombined in FP64 with the power-of-two
// scales.  The K = nocc half and the energy epilogue stay FP64.

// BF16x3 form: the row is scaled by 2^-e(x) so its largest entry is below 1,
// then written as three BF16 planes h1 + h2 + h3 (h1 = RN(v), h2 = RN(v - h1),
// h3 = RN(v - h1 - h2); the residual is below 2^-24 |v|).
__device__ __forceinline__ double bf16_rn(double v, uint16_t &bits)
{
    // to FP32 by round-to-nearest, then to BF16 by round-to-nearest-even (the
    // residual of the double rounding lands in the next plane)
    const uint32_t u = __builtin_bit_cast(uint32_t, (float)v);
    const uint32_t r = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    bits = (uint16_t)(r >> 16);
    return (double)__builtin_bit_cast(float, r);
}

// Split-form packing, one workgroup (4 waves) per 16-row block of one panel.
// Plane layout per 16-row block xb and K step ks of the MFMA (KS = 64 INT8
// digits or 32 BF16 values): lane l holds 16 bytes for row x = 16*xb + l%16 at
// K offset ks*KS + (KS/4)*(l/16) -- the operand layout of
// v_mfma_i32_16x16x64_i8 and v_mfma_f32_16x16x32_bf16 -- so one K step of one
// plane is 1 KiB.  Thread t of a step handles row r = (t/4)%16, K group
// (t/4)/16 (one per wave) and a quarter of that lane's 16 bytes: its reads of
// P are 128-byte runs whichever way P is transposed, and its 4-byte stores
// are contiguous across the workgroup.  Each part of the row is split in two
// passes: pass 1 takes the row's maximum (lane shuffles within a wave, then LDS
// across the 4 waves), pass 2 writes the planes.  The K = nvir part (P) fills
// steps 0..nkv-1 of the block.  T8 (INT8 with occupied digit planes): the K =
// nocc part (sgn * Q) fills steps nkv..nkv+nkx-1 with its own row scale (second
// scale set); otherwise it is written as an FP64 tail panel.
template <int S, bool BF, bool T8 = false>
__global__ __launch_bounds__(256) void k_split_pack(const PanelDesc *__restrict__ descs, Geom g,
                                                   SplitGeom sg, int8_t *__restrict__ planes,
                                                   double *__restrict__ scales,
                                                   double *__restrict__ tail)
{
    constexpr int KS = BF ? 32 : 64;     // K per MFMA step
    constexpr int EPT = BF ? 2 : 4;      // values per thread per step (4 bytes per plane)
    constexpr int NPL = BF ? 3 : S;      // planes
    static_assert(!T8 || (!BF && S <= kTail8MaxSlices), "occupied digit planes: INT8, S <= 3");
    __shared__ double rmax[4][16];
    const int pnl = blockIdx.y, xb = blockIdx.x, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int r = (lane >> 2) & 15, c = lane & 3;
    const int x = xb * 16 + r;
    const PanelDesc d = descs[pnl];
    const int nv = g.nvir, no = g.nocc;
    // K index of value j of this thread in step ks of a part
    auto kidx = [&](int ks, int j) { return ks * KS + w * (KS / 4) + c * EPT + j; };
    uint32_t *pl = reinterpret_cast<uint32_t *>(planes);
    const int64_t plane_words = sg.slice_bytes / 4;
    const int nsb = sg.nkv + sg.nkx;  // K steps of one block

    // one part of the row: val(k) for k in [0, nsteps * KS), steps step0.. of the block
    auto split_part = [&](auto val, int nsteps, int step0, double *scale_set) {
        double m = 0.0;
        for (int ks = 0; ks < nsteps; ++ks)
#pragma unroll
            for (int j = 0; j < EPT; ++j) m = fmax(m, fabs(val(kidx(ks, j))));
        m = fmax(m, __shfl_xor(m, 1, 64));
        m = fmax(m, __shfl_xor(m, 2, 64));
        if (c == 0) rmax[w][r] = m;
        __syncthreads();
        m = fmax(fmax(rmax[0][r], rmax[1][r]), fmax(rmax[2][r], rmax[3][r]));
        int e = 0;
        // INT8: largest entry at most 127/128 after scaling; BF16: below 1
        if (m > 0.0) (void)frexp(BF ? m : m * (128.0 / 127.0), &e);
        const double inv = ldexp(1.0, -e);
        if (w == 0 && c == 0) scale_set[(int64_t)pnl * g.nvp + x] = ldexp(1.0, e);
        for (int ks = 0; ks < nsteps; ++ks) {
            uint32_t word[NPL];
#pragma unroll
            for (int q = 0; q < NPL; ++q) word[q] = 0;
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                double v = val(kidx(ks, j)) * inv;
#pragma unroll
                for (int q = 0; q < NPL; ++q) {
                    if (BF) {
                        uint16_t b;
                        v -= bf16_rn(v, b);
                        word[q] |= (uint32_t)b << (16 * j);
                    } else {
                        v *= 128.0;
                        const double dg = rint(v);
                        v -= dg;
                        word[q] |= (uint32_t)(uint8_t)(int8_t)dg << (8 * j);
                    }
                }
            }
            // byte offset of this thread's 4 bytes: step base + lane-group (w*16 + r)
            // * 16 + c * 4
            const int64_t o = ((int64_t)(xb * nsb + step0 + ks) * 1024 + (w * 16 + r) * 16 + c * 4) / 4;
#pragma unroll
            for (int q = 0; q < NPL; ++q) pl[((int64_t)pnl * NPL + q) * plane_words + o] = word[q];
        }
        __syncthreads();  // rmax is reused by the next part
    };
    split_part([&](int k) -> double {
        if (x >= nv || k >= nv) return 0.0;
        return d.first_t ? d.P[k + (int64_t)x * nv] : d.P[x + (int64_t)k * nv];
    }, sg.nkv, 0, scales);
    if (T8) {
        auto qv = [&](int k) -> double {
            if (x >= nv || k >= no) return 0.0;
            return d.sgn * (d.second_t ? d.Q[k + (int64_t)x * no] : d.Q[x + (int64_t)k * nv]);
        };
        double *sset = scales + (int64_t)sg.npanels * g.nvp;
        if (!sg.fold) {
            split_part(qv, sg.nkx, sg.nkv, sset);
            return;
        }
        // Folded (nocc <= 32): a 64-wide K step holds two digits of the same
        // element k < 32, one in K 0..31 and one in K 32..63, so a product of two
        // such slots sums two digit products of one level.  A panel is an A
        // operand in one wave and a B operand in the other, so both forms are
        // stored: A: [d1|d2] (plane 0, step nkv), [d3|0] (plane 1, step nkv);
        // B: [d1|0] (plane 2, step nkv), [d2|d1] (plane 0, step nkv+1), [d3|d2]
        // (plane 1, step nkv+1); plane 2 of step nkv+1 is zero.  Then level 0 =
        // [d1|d2].[d1|0], level 1 = [d1|d2].[d2|d1], level 2 = [d1|d2].[d3|d2] +
        // [d3|0].[d1|0].  The digits are those the unfolded form computes.
        const bool lo = w < 2;
        const int kb = (w & 1) * 16 + c * 4;  // this thread's four elements
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmax(m, fabs(qv(kb + j)));
        m = fmax(m, __shfl_xor(m, 1, 64));
        m = fmax(m, __shfl_xor(m, 2, 64));
        if (c == 0) rmax[w][r] = m;
        __syncthreads();
        m = fmax(fmax(rmax[0][r], rmax[1][r]), fmax(rmax[2][r], rmax[3][r]));
        int e = 0;
        if (m > 0.0) (void)frexp(m * (128.0 / 127.0), &e);
        const double inv = ldexp(1.0, -e);
        if (w == 0 && c == 0) sset[(int64_t)pnl * g.nvp + x] = ldexp(1.0, e);
        uint32_t word[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double v = qv(kb + j) * inv;
            int dg[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                v *= 128.0;
                const double t = rint(v);
                v -= t;
                dg[q] = (int)t;
            }
            const int slot[6] = {lo ? dg[0] : dg[1], lo ? dg[2] : 0, lo ? dg[0] : 0,
                                 lo ? dg[1] : dg[0], lo ? dg[2] : dg[1], 0};
#pragma unroll
            for (int q = 0; q < 6; ++q) word[q] |= (uint32_t)(uint8_t)(int8_t)slot[q] << (8 * j);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int plane_q = q % 3, step = sg.nkv + q / 3;
            const int64_t o = ((int64_t)(xb * nsb + step) * 1024 + (w * 16 + r) * 16 + c * 4) / 4;
            pl[((int64_t)pnl * NPL + plane_q) * plane_words + o] = word[q];
        }
        return;
    }
    // FP64 tail in the step-pair layout (nkt is even): element 128*s + 2*l + h
    // holds row l%16, K index 4*(2s+h) + l/16
    double *tp = tail + (int64_t)pnl * sg.tail_sz + (int64_t)xb * sg.nkt * 64;
    for (int e2 = t; e2 < sg.nkt * 64; e2 += 256) {
        const int l = (e2 >> 1) & 63;
        const int kk = 4 * (2 * (e2 >> 7) + (e2 & 1)) + (l >> 4);
        const int xx = xb * 16 + (l & 15);
        double v = 0.0;
        if (xx < nv && kk < no)
            v = d.sgn * (d.second_t ? d.Q[kk + (int64_t)xx * no] : d.Q[xx + (int64_t)kk * nv]);
        tp[e2] = v;
    }
}

hipError_t launch_split_pack(const PanelDesc *d_desc, int npanels, const Geom &g,
                             const SplitGeom &sg, int8_t *digits, double *scales, double *tail,
                             hipStream_t st)
{
    dim3 grid(g.nb, npanels), block(256);
    if (sg.bf16) {
        hipLaunchKernelGGL((k_split_pack<3, true>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail);
        return hipGetLastError();
    }
    if (sg.nkx > 0) {
        if (sg.slices != 3) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_split_pack<3, false, true>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail);
        return hipGetLastError();
    }
    switch (sg.slices) {
    case 3: hipLaunchKernelGGL((k_split_pack<3, false>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail); break;
    case 4: hipLaunchKernelGGL((k_split_pack<4, false>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail); break;
    case 5: hipLaunchKernelGGL((k_split_pack<5, false>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail); break;
    case 6: hipLaunchKernelGGL((k_split_pack<6, false>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail); break;
    case 7: hipLaunchKernelGGL((k_split_pack<7, false>), grid, block, 0, st, d_desc, g, sg, digits, scales, tail); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

typedef int int4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int4v buffer_load_i8x16(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(int4v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// Work unit of k_fused (one 16x16 block pair of one triple) split over two
// waves: wave 0 computes the eight products at (b,c) (groups n and t), wave 1 the
// eight mirrored ones f(c,b) (groups n' and t').  Each group shares two A and two
// B panels.  After the INT8 groups each wave adds its FP64 tail (K = nocc) on
// v_mfma_f64_16x16x4_f64 with the A rows fed in the order 4*(i%4) + i/4, which
// makes its result layout (lane l, element r -> row 4*(l/16)+r, column l%16)
// equal that of the INT8 MFMA; with occupied digit planes (T8) each group
// instead runs the K = nocc steps as INT8 digit products too.  The energy epilogue swaps half the values
// through LDS and each wave finishes two of the four rows per lane.
// The eight products of one wave role (W = 0: (b,c); W = 1: mirrored (c,b)),
// panel ids resolved at compile time so every operand offset stays scalar.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Digit-plane arithmetic of the two split forms: INT8 digits on
// v_mfma_i32_16x16x64_i8 with INT32 accumulators, or BF16 planes on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulators.  Both take 16 bytes per lane
// per operand and leave a 16x16 tile in the layout row 4*(l/16)+r, column l%16.
template <bool BF>
struct PlaneMath;

template <>
struct PlaneMath<false> {
    typedef int4v acc_t;
    static __device__ __forceinline__ acc_t mfma(int4v a, int4v b, acc_t c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
    // weight of accumulation level L (= plane p + plane q, 0-based): 2^(-7(L+2))
    static __device__ __forceinline__ double weight(int L) { return __builtin_ldexp(1.0, -7 * (L + 2)); }
};

template <>
struct PlaneMath<true> {
    typedef flt4 acc_t;
    static __device__ __forceinline__ acc_t mfma(int4v a, int4v b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
    static __device__ __forceinline__ double weight(int) { return 1.0; }
};

// One lane's row r of a digit-product tile: V = sum_L acc_L 2^(7 (S-1-L)), an
// integer below 2^53 and so exact in FP64, equal to sum_L acc_L w(L) / w(S-1).
// Adjacent levels are first combined in INT32, (acc_L << 7) + acc_(L+1), exact
// when every level sum has at most kIntPairMaxK terms (S <= 4): one conversion
// and one FMA fewer per pair than converting every level.
template <int S>
__device__ __forceinline__ double level_sum(const int4v (&acc)[S], int r)
{
    double u;
    int q;
    if constexpr (S % 2 == 1) {
        u = (double)acc[0][r];
        q = 1;
    } else {
        u = (double)((acc[0][r] << 7) + acc[1][r]);
        q = 2;
    }
#pragma unroll
    for (; q + 1 < S; q += 2) u = __builtin_fma(u, 16384.0, (double)((acc[q][r] << 7) + acc[q + 1][r]));
    return u;
}

// The same integer V for any K: every level converted, then Horner steps by 2^7
// (each partial value an integer below 2^53, so every step is exact).
template <int S>
__device__ __forceinline__ double level_horner(const int4v (&acc)[S], int r)
{
    double u = (double)acc[0][r];
#pragma unroll
    for (int q = 1; q < S; ++q) u = __builtin_fma(u, 128.0, (double)acc[q][r]);
    return u;
}

// stash: this wave's 8 KiB of LDS, holding group 0's four products while group
// 1's digit products run, so their 32 VGPRs are free during the INT8 loops.
// sst: this wave's 2 KiB of LDS for the row scales of its eight panels (p = A0,
// A1 of group 0, A0, A1 of group 1, then B0, B1 of each group; 16 rows each of
// block bi for A, ci for B), fetched with one 16-byte load per lane: the K =
// nvir planes' scales, then (T8) the K = nocc planes' scales.
// T8 (INT8 with occupied digit planes): each group's K = nvir steps are
// recombined first, then its K = nocc steps (own scales) are added in FP64.
// Otherwise the K = nocc half runs as an FP64 tail on v_mfma_f64_16x16x4_f64
// after both groups.
template <int S, int W, bool BF = false, bool T8 = false>
__device__ __forceinline__ void split_wave(dbl4 (&facc)[8], const __amdgpu_buffer_rsrc_t rd,
                                           const __amdgpu_buffer_rsrc_t rt,
                                           const double *__restrict__ scales, const Geom &g,
                                           const SplitGeom &sg, const PanelPairs &pp, int bi,
                                           int ci, int lane, double *stash,
                                           double *__restrict__ sst)
{
    // panel ids: pair q holds panels 2q, 2q+1
    const int LJ = 2 * pp.li, LK = LJ + 1, LJT = 2 * pp.lk, LKT = LJT + 1;
    const int R1 = 2 * pp.rkj, R3 = R1 + 1, R1T = 2 * pp.rij, R3T = R1T + 1;
    const int nvp = g.nvp;
    static_assert(!T8 || (!BF && S <= kTail8MaxSlices), "occupied digit planes: INT8, S <= 3");
    dbl2 sreg, sreg_t = {0.0, 0.0};
    {
        const int sp = lane >> 3;
        const int L4[4] = {LJ, LK, LJT, LKT}, R4[4] = {R1, R3, R1T, R3T};
        int pid = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pid = sp == q ? (W == 0 ? L4[q] : R4[q]) : pid;
            pid = sp == 4 + q ? (W == 0 ? R4[q] : L4[q]) : pid;
        }
        const int64_t so = (int64_t)pid * nvp + (sp < 4 ? bi : ci) * 16 + (lane & 7) * 2;
        sreg = gload2(scales + so);
        if (T8) sreg_t = gload2(scales + (int64_t)sg.npanels * nvp + so);
    }
    const uint32_t plane = (uint32_t)sg.slice_bytes;
    const uint32_t blk8 = (uint32_t)(sg.nkv + sg.nkx) * 1024;  // bytes of a 16-row block of one plane
    const uint32_t voff8 = lane * 16;
    const int row0 = 4 * (lane >> 4), cx = ci * 16 + (lane & 15);
    // INT8: weight of the lowest level, 2^(-7(S+1)), applied with the column
    // scale; adjacent levels pair in INT32 when no level sum exceeds
    // kIntPairMaxK terms (S <= 4)
    const double wlow = BF ? 1.0 : __builtin_ldexp(1.0, -7 * (S + 1));
    const bool pairable = !BF && S <= 4 && sg.nkv * 64 <= kIntPairMaxK &&
                          (!T8 || sg.fold || sg.nkx * 64 <= kIntPairMaxK);
#pragma unroll
    for (int gi = 0; gi < 2; ++gi) {
        // panels {A0, A1, B0, B1}: W=0 groups n, t; W=1 groups n', t'
        int A0, A1, B0, B1;
        if (W == 0) {
            A0 = gi == 0 ? LJ : LJT;  A1 = gi == 0 ? LK : LKT;
            B0 = gi == 0 ? R1 : R1T;  B1 = gi == 0 ? R3 : R3T;
        } else {
            A0 = gi == 0 ? R1 : R1T;  A1 = gi == 0 ? R3 : R3T;
            B0 = gi == 0 ? LJ : LJT;  B1 = gi == 0 ? LK : LKT;
        }
        typedef typename PlaneMath<BF>::acc_t acc_t;
        const acc_t zacc = {0, 0, 0, 0};
        acc_t iacc[4][S];
        // (cleared here: peeling the first K step to start each level from a
        // zero operand instead raised the kernel above 168 VGPRs)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < S; ++q) iacc[c][q] = zacc;
        const uint32_t oa0 = (uint32_t)A0 * S * plane + bi * blk8;
        const uint32_t oa1 = (uint32_t)A1 * S * plane + bi * blk8;
        const uint32_t ob0 = (uint32_t)B0 * S * plane + ci * blk8;
        const uint32_t ob1 = (uint32_t)B1 * S * plane + ci * blk8;
        // ping-pong operand buffers X and Y, two K steps per trip: each buffer's
        // planes are in flight while the other buffer's digit products run; an
        // odd last step runs from X after the loop (steps stay in order)
        int4v xa0[S], xa1[S], xb0[S], xb1[S], ya0[S], ya1[S], yb0[S], yb1[S];
#define SPLIT_LOAD(A0_, A1_, B0_, B1_, kstep)                                                    \
        _Pragma("unroll") for (int q = 0; q < S; ++q) {                                         \
            const uint32_t ko_ = (uint32_t)(kstep) * 1024;                                        \
            A0_[q] = buffer_load_i8x16(rd, voff8, oa0 + q * plane + ko_);                        \
            A1_[q] = buffer_load_i8x16(rd, voff8, oa1 + q * plane + ko_);                        \
            B0_[q] = buffer_load_i8x16(rd, voff8, ob0 + q * plane + ko_);                        \
            B1_[q] = buffer_load_i8x16(rd, voff8, ob1 + q * plane + ko_);                        \
        }
#define SPLIT_MFMA_Z(A0_, A1_, B0_, B1_, first)                                                  \
        _Pragma("unroll") for (int sa = 0; sa < S; ++sa) {                                      \
            _Pragma("unroll") for (int sb = 0; sb < S - sa; ++sb) {                             \
                const int lv = sa + sb;                                                          \
                const bool z_ = (first) && sa == 0;                                              \
                iacc[0][lv] = PlaneMath<BF>::mfma(A0_[sa], B0_[sb], z_ ? zacc : iacc[0][lv]);     \
                iacc[1][lv] = PlaneMath<BF>::mfma(A1_[sa], B0_[sb], z_ ? zacc : iacc[1][lv]);     \
                iacc[2][lv] = PlaneMath<BF>::mfma(A0_[sa], B1_[sb], z_ ? zacc : iacc[2][lv]);     \
                iacc[3][lv] = PlaneMath<BF>::mfma(A1_[sa], B1_[sb], z_ ? zacc : iacc[3][lv]);     \
            }                                                                                    \
        }
#define SPLIT_MFMA(A0_, A1_, B0_, B1_) SPLIT_MFMA_Z(A0_, A1_, B0_, B1_, false)
        const int nkv = sg.nkv;
        if constexpr (S <= 4 && !BF) {
            // step 0 is peeled: its level-opening products take a zero
            // accumulator, so the accumulators need no clearing
            SPLIT_LOAD(xa0, xa1, xb0, xb1, 0)
            int ks = 0;
            for (; ks + 1 < nkv; ks += 2) {
                SPLIT_LOAD(ya0, ya1, yb0, yb1, ks + 1)
                SPLIT_MFMA(xa0, xa1, xb0, xb1)
                const int k2 = (ks + 2 < nkv) ? ks + 2 : nkv - 1;  // clamped prefetch, in bounds
                SPLIT_LOAD(xa0, xa1, xb0, xb1, k2)
                SPLIT_MFMA(ya0, ya1, yb0, yb1)
            }
            if (nkv & 1) SPLIT_MFMA(xa0, xa1, xb0, xb1)
        } else {
            // five or more planes: a second buffer would cost a wave per SIMD
            // (the BF16 planes measured slower with it)
            for (int ks = 0; ks < nkv; ++ks) {
                SPLIT_LOAD(xa0, xa1, xb0, xb1, ks)
                SPLIT_MFMA(xa0, xa1, xb0, xb1)
            }
        }
        // recombine: f = 2^(eA(b) + eB(c)) * sum_L acc_L * w(L) (INT8 w(L) =
        // 2^(-7(L+2)), BF16 w = 1); the combination (a, b) is product
        // 4*gi + a + 2*b (W=0) or 4*gi + 2*a + b (W=1).  With T8 a second pass
        // adds the K = nocc steps with their own scales (sst + 128).  INT8: the
        // levels are summed as the exact integer V = sum_L acc_L 2^(7(S-1-L))
        // (level_sum when the levels pair in INT32, else level_horner) and
        // f = V sA (sB 2^(-7(S+1))): every step is exact (powers of two and
        // integers below 2^53), so f is bitwise the weighted sum's.
        auto recombine = [&](const double *ss, bool add, auto pairc) {
            constexpr bool PAIR = decltype(pairc)::value;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int ca = c & 1, cb = c >> 1;
                const int lp = 4 * gi + (W == 0 ? ca + 2 * cb : 2 * ca + cb);
                const double sB = ss[(4 + 2 * gi + cb) * 16 + (lane & 15)] * wlow;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double v;
                    if constexpr (BF) {
                        v = 0.0;
#pragma unroll
                        for (int q = S - 1; q >= 0; --q) v += (double)iacc[c][q][r];
                    } else if constexpr (PAIR) {
                        v = level_sum<S>(iacc[c], r);
                    } else {
                        v = level_horner<S>(iacc[c], r);
                    }
                    const double sA = ss[(2 * gi + ca) * 16 + row0 + r];
                    const double f = v * sA * sB;
                    if (gi == 0) {
                        double &st = stash[(lp * 4 + r) * 64 + lane];  // lp = 0..3
                        st = add ? st + f : f;
                    } else {
                        facc[lp][r] = add ? facc[lp][r] + f : f;
                    }
                }
            }
        };
        if (gi == 0) *reinterpret_cast<dbl2 *>(sst + 2 * lane) = sreg;
        if (pairable) recombine(sst, false, std::true_type{});
        else recombine(sst, false, std::false_type{});
        // a compiler fence (stash is not restrict-qualified, so it covers the
        // stash): without it the stores are forwarded to the reloads below and
        // group 0's values stay in registers through group 1 after all
        if (gi == 0) asm volatile("" ::: "memory");
        if constexpr (T8) {
            if (gi == 0) *reinterpret_cast<dbl2 *>(sst + 128 + 2 * lane) = sreg_t;
            // (the scheduling barrier keeps these loads from being hoisted above the
            // recombination, where their registers would add to its own)
            __builtin_amdgcn_sched_barrier(0);
            if (sg.fold) {
                // folded slots (k_split_pack, nocc <= 32): A [d1|d2], [d3|0] and
                // B [d1|0], [d2|d1], [d3|d2]; four MFMAs per product for the
                // three levels, 5 loads per operand pair instead of 6
                const uint32_t k0 = (uint32_t)nkv * 1024, k1 = k0 + 1024;
                xa0[0] = buffer_load_i8x16(rd, voff8, oa0 + k0);
                xa1[0] = buffer_load_i8x16(rd, voff8, oa1 + k0);
                xa0[1] = buffer_load_i8x16(rd, voff8, oa0 + plane + k0);
                xa1[1] = buffer_load_i8x16(rd, voff8, oa1 + plane + k0);
                xb0[0] = buffer_load_i8x16(rd, voff8, ob0 + 2 * plane + k0);
                xb1[0] = buffer_load_i8x16(rd, voff8, ob1 + 2 * plane + k0);
                xb0[1] = buffer_load_i8x16(rd, voff8, ob0 + k1);
                xb1[1] = buffer_load_i8x16(rd, voff8, ob1 + k1);
                xb0[2] = buffer_load_i8x16(rd, voff8, ob0 + plane + k1);
                xb1[2] = buffer_load_i8x16(rd, voff8, ob1 + plane + k1);
#pragma unroll
                for (int lv = 0; lv < 3; ++lv) {
                    iacc[0][lv] = PlaneMath<BF>::mfma(xa0[0], xb0[lv], zacc);
                    iacc[1][lv] = PlaneMath<BF>::mfma(xa1[0], xb0[lv], zacc);
                    iacc[2][lv] = PlaneMath<BF>::mfma(xa0[0], xb1[lv], zacc);
                    iacc[3][lv] = PlaneMath<BF>::mfma(xa1[0], xb1[lv], zacc);
                }
                iacc[0][2] = PlaneMath<BF>::mfma(xa0[1], xb0[0], iacc[0][2]);
                iacc[1][2] = PlaneMath<BF>::mfma(xa1[1], xb0[0], iacc[1][2]);
                iacc[2][2] = PlaneMath<BF>::mfma(xa0[1], xb1[0], iacc[2][2]);
                iacc[3][2] = PlaneMath<BF>::mfma(xa1[1], xb1[0], iacc[3][2]);
            } else {
                SPLIT_LOAD(xa0, xa1, xb0, xb1, nkv)
                SPLIT_MFMA_Z(xa0, xa1, xb0, xb1, true)
                for (int kx = 1; kx < sg.nkx; ++kx) {
                    SPLIT_LOAD(xa0, xa1, xb0, xb1, nkv + kx)
                    SPLIT_MFMA(xa0, xa1, xb0, xb1)
                }
            }
            if (pairable) recombine(sst + 128, true, std::true_type{});
            else recombine(sst + 128, true, std::false_type{});
            if (gi == 0) asm volatile("" ::: "memory");
            // group 1's first loads stay below this group's last recombination
            __builtin_amdgcn_sched_barrier(0);
        }
#undef SPLIT_LOAD
#undef SPLIT_MFMA
#undef SPLIT_MFMA_Z
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) facc[p][r] = stash[(p * 4 + r) * 64 + lane];

    // FP64 tail (K = nocc) of the eight products
    if constexpr (!T8) {
        const uint32_t tsb = (uint32_t)(sg.tail_sz * 8);
        const uint32_t blkt = (uint32_t)sg.nkt * 512;
        const int li = lane & 15;
        // step-pair layout: each lane's two doubles of a step pair are adjacent
        const uint32_t voffA = (uint32_t)((4 * (li & 3) + (li >> 2)) + (lane & 48)) * 16;
        const uint32_t voffB = lane * 16;
        // W=0: A {LJ, LK, LJt, LKt} at b, B {R1, R3, R1t, R3t} at c; W=1 roles swapped
        const int PL[4] = {LJ, LK, LJT, LKT};
        const int PR[4] = {R1, R3, R1T, R3T};
        uint32_t ao[4], bo[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ao[v] = (W == 0 ? PL[v] : PR[v]) * tsb + bi * blkt;
            bo[v] = (W == 0 ? PR[v] : PL[v]) * tsb + ci * blkt;
        }
        for (int sp = 0; sp < sg.nkt / 2; ++sp) {
            dbl2 a2[4], b2[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                a2[v] = buffer_load_f64x2(rt, voffA, ao[v] + sp * 1024);
                b2[v] = buffer_load_f64x2(rt, voffB, bo[v] + sp * 1024);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                double a[4], b[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    a[v] = a2[v][h];
                    b[v] = b2[v][h];
                }
#define MS(pp, ia, ib) facc[pp] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ia], b[ib], facc[pp], 0, 0, 0)
                if (W == 0) {
                    MS(0, 0, 0); MS(1, 1, 0); MS(2, 0, 1); MS(3, 1, 1);
                    MS(4, 2, 2); MS(5, 3, 2); MS(6, 2, 3); MS(7, 3, 3);
                } else {
                    MS(0, 0, 0); MS(1, 0, 1); MS(2, 1, 0); MS(3, 1, 1);
                    MS(4, 2, 2); MS(5, 2, 3); MS(6, 3, 2); MS(7, 3, 3);
                }
#undef MS
            }
        }
    }
}

// FP64 form of the two-wave item (k_fused2): the same eight products per wave
// role as split_wave, over the full K = nvir + nocc of the FP64 panels in the
// step-pair layout.  Each pair of K steps takes four A and four B 16-byte loads
// and runs sixteen v_mfma_f64_16x16x4_f64, in step order; ping-pong buffers of
// step pairs.  The A rows are fed in the order 4*(i%4) + i/4 so the result
// layout is the one pair_epilogue expects.
template <int W>
__device__ __forceinline__ void fp64_wave(dbl4 (&facc)[8], const double *__restrict__ pan,
                                          const Geom &g, const PanelPairs &pp, int bi, int ci,
                                          int lane)
{
    __amdgpu_buffer_rsrc_t rp[4];
    pair_resources(rp, pan, g.vecsz, pp);
    const uint32_t vsb = (uint32_t)(g.vecsz * 8);
    const uint32_t blkb = (uint32_t)g.nks * 512;
    const uint32_t ob = bi * blkb, oc = ci * blkb;
    const int li = lane & 15;
    const uint32_t la = (uint32_t)((4 * (li & 3) + (li >> 2)) + (lane & 48));
    // panels in the step-pair layout (k_pack<true>): 16 bytes per lane per pair
    const uint32_t voffA2 = la * 16, voffB2 = lane * 16;
    // L panels {LJ, LK, LJt, LKt} and R panels {R1, R3, R1t, R3t}: their pair,
    // and first (v even) or second panel of it
    constexpr int LR[4] = {RP_LI, RP_LI, RP_LK, RP_LK};
    constexpr int RR[4] = {RP_RKJ, RP_RKJ, RP_RIJ, RP_RIJ};
    uint32_t ao[4], bo[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ao[v] = (v & 1) * vsb + ob;
        bo[v] = (v & 1) * vsb + oc;
    }
    const int nks = g.nks, npair = nks >> 1;
    dbl2 xa[4], xb[4], ya[4], yb[4];
#define F2_LOAD(da, db, spair)                                                                   \
    _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                             \
        const int ra = (W == 0) ? LR[v] : RR[v], rb = (W == 0) ? RR[v] : LR[v];                   \
        da[v] = buffer_load_f64x2(rp[ra], voffA2, ao[v] + (uint32_t)(spair) * 1024);             \
        db[v] = buffer_load_f64x2(rp[rb], voffB2, bo[v] + (uint32_t)(spair) * 1024);             \
    }
#define MS(pp, ia, ib) facc[pp] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ia], b[ib], facc[pp], 0, 0, 0)
#define F2_STEP(a, b)                                                                            \
    {                                                                                            \
        if (W == 0) {                                                                            \
            MS(0, 0, 0); MS(1, 1, 0); MS(2, 0, 1); MS(3, 1, 1);                                  \
            MS(4, 2, 2); MS(5, 3, 2); MS(6, 2, 3); MS(7, 3, 3);                                  \
        } else {                                                                                 \
            MS(0, 0, 0); MS(1, 0, 1); MS(2, 1, 0); MS(3, 1, 1);                                  \
            MS(4, 2, 2); MS(5, 2, 3); MS(6, 3, 2); MS(7, 3, 3);                                  \
        }                                                                                        \
    }
#define F2_MFMA(da, db)                                                                          \
    _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                             \
        double a[4], b[4];                                                                       \
        _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                         \
            a[v] = da[v][h];                                                                     \
            b[v] = db[v][h];                                                                     \
        }                                                                                        \
        F2_STEP(a, b)                                                                            \
    }
    if (npair > 0) F2_LOAD(xa, xb, 0)
    int sp = 0;
    for (; sp + 1 < npair; sp += 2) {
        F2_LOAD(ya, yb, sp + 1)
        F2_MFMA(xa, xb)
        const int s2 = (sp + 2 < npair) ? sp + 2 : npair - 1;  // clamped prefetch, in bounds
        F2_LOAD(xa, xb, s2)
        F2_MFMA(ya, yb)
    }
    if (npair & 1) F2_MFMA(xa, xb)
    if (nks & 1) {
        // odd last step, single-step order
        double a[4], b[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ra = (W == 0) ? LR[v] : RR[v], rb = (W == 0) ? RR[v] : LR[v];
            a[v] = buffer_load_f64(rp[ra], la * 8, ao[v] + (uint32_t)npair * 1024);
            b[v] = buffer_load_f64(rp[rb], lane * 8, bo[v] + (uint32_t)npair * 1024);
        }
        F2_STEP(a, b)
    }
#undef F2_LOAD
#undef F2_MFMA
#undef F2_STEP
#undef MS
}

// FP32 form of the two-wave item (k_fused32): fp64_wave's eight products per
// wave role over FP32 panels on v_mfma_f32_16x16x4_f32, FP32 accumulation.
// That MFMA's result layout is already row 4*(l/16)+r, column l%16, so the A
// rows go in natural order.
__device__ __forceinline__ float buffer_load_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

template <int W>
__device__ __forceinline__ void fp32_wave(flt4 (&facc)[8], const float *__restrict__ pan,
                                          const Geom &g, const PanelPairs &pp, int bi, int ci,
                                          int lane)
{
    const int64_t vz = g.vecsz;
    const uint32_t bytes = (uint32_t)(2 * vz * 4);
    __amdgpu_buffer_rsrc_t rp[4];
    rp[RP_LI] = buffer_rsrc(reinterpret_cast<const double *>(pan + 2 * (int64_t)pp.li * vz), bytes);
    rp[RP_LK] = buffer_rsrc(reinterpret_cast<const double *>(pan + 2 * (int64_t)pp.lk * vz), bytes);
    rp[RP_RIJ] = buffer_rsrc(reinterpret_cast<const double *>(pan + 2 * (int64_t)pp.rij * vz), bytes);
    rp[RP_RKJ] = buffer_rsrc(reinterpret_cast<const double *>(pan + 2 * (int64_t)pp.rkj * vz), bytes);
    const uint32_t vsb = (uint32_t)(vz * 4);
    const uint32_t blkb = (uint32_t)g.nks * 256;
    const uint32_t ob = bi * blkb, oc = ci * blkb;
    const uint32_t voff = lane * 4;
    constexpr int LR[4] = {RP_LI, RP_LI, RP_LK, RP_LK};
    constexpr int RR[4] = {RP_RKJ, RP_RKJ, RP_RIJ, RP_RIJ};
    uint32_t ao[4], bo[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ao[v] = (v & 1) * vsb + ob;
        bo[v] = (v & 1) * vsb + oc;
    }
    const int nks = g.nks;
    float xa[4], xb[4], ya[4], yb[4];
#define F3_LOAD(da, db, kstep)                                                                   \
    _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                             \
        const int ra = (W == 0) ? LR[v] : RR[v], rb = (W == 0) ? RR[v] : LR[v];                   \
        da[v] = buffer_load_f32(rp[ra], voff, ao[v] + (uint32_t)(kstep) * 256);                   \
        db[v] = buffer_load_f32(rp[rb], voff, bo[v] + (uint32_t)(kstep) * 256);                   \
    }
#define MS(pp_, ia, ib) facc[pp_] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ia], b[ib], facc[pp_], 0, 0, 0)
#define F3_MFMA(da, db)                                                                          \
    {                                                                                            \
        const float *a = da, *b = db;                                                            \
        if (W == 0) {                                                                            \
            MS(0, 0, 0); MS(1, 1, 0); MS(2, 0, 1); MS(3, 1, 1);                                  \
            MS(4, 2, 2); MS(5, 3, 2); MS(6, 2, 3); MS(7, 3, 3);                                  \
        } else {                                                                                 \
            MS(0, 0, 0); MS(1, 0, 1); MS(2, 1, 0); MS(3, 1, 1);                                  \
            MS(4, 2, 2); MS(5, 2, 3); MS(6, 3, 2); MS(7, 3, 3);                                  \
        }                                                                                        \
    }
    F3_LOAD(xa, xb, 0)
    int ks = 0;
    for (; ks + 1 < nks; ks += 2) {
        F3_LOAD(ya, yb, ks + 1)
        F3_MFMA(xa, xb)
        const int k2 = (ks + 2 < nks) ? ks + 2 : nks - 1;  // clamped prefetch, in bounds
        F3_LOAD(xa, xb, k2)
        F3_MFMA(ya, yb)
    }
    if (nks & 1) F3_MFMA(xa, xb)
#undef F3_LOAD
#undef F3_MFMA
#undef MS
}

// Epilogue of a two-wave block-pair item.  On entry wave 0 holds the eight
// products at (b,c) and wave 1 the eight mirrored ones, both in the layout
// lane l, element r -> row 4*(l/16)+r, column l%16.  On a diagonal pair
// (B = C) wave 1 computed nothing: the mirrored products are the direct ones
// transposed, read back from one LDS tile.
// T = float (the FP32 form) evaluates the energy summand in FP32, with the
// denominator formed as in gpu-omp/ccsd_trpdrv_omp_cbody.c:178 (FP32 sum, FP64
// division, rounded to FP32); the per-pair partial leaves as FP64.
template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef dbl4 type; };
template <> struct Vec4<float> { typedef flt4 type; };

// The energy vectors of the item's two blocks: vst[q][s][i] holds eorb (q = 0)
// or vector q-1 (t1v1 t1v2 dintx1 dintx2 dintc1 dintc2) at row s ? c : b block
// * 16 + i.  Each wave fetches half the table with one 16-byte load per lane at
// the start of the item (stage_vectors); pair_epilogue writes it to LDS.
__device__ __forceinline__ dbl2 stage_vectors(const JobDev &jd, const Task &tk, int bi, int ci, int w,
                                              int lane)
{
    dbl2 v = {0.0, 0.0};
    const int e = w * 112 + 2 * lane;  // element pair of this lane (lanes 0..55)
    if (lane < 56) {
        const int q = e >> 5, side = (e >> 4) & 1, i = e & 15;
        int row = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r) row = q == r + 1 ? uniform(tk.vrow[r]) : row;
        const double *src = q == 0 ? jd.ev : jd.vec6 + (int64_t)row * jd.g.nvp;
        v = gload2(src + (side ? ci : bi) * 16 + i);
    }
    return v;
}

template <typename T, bool TENGY, bool WRITEF>
__device__ __forceinline__ void pair_epilogue(const JobDev &jd, const Task &tk,
                                              const typename Vec4<T>::type (&facc)[8], int bi, int ci,
                                              int w, int lane, double *__restrict__ part4,
                                              double *__restrict__ fo, void *lds, dbl2 vreg,
                                              double *__restrict__ vst, bool ftiled)
{
    // the workgroup's 16 KiB exchange buffer (each wave's half first held its
    // stash, consumed before the products were finished)
    T(*xch)[8][2][64] = reinterpret_cast<T(*)[8][2][64]>(lds);
    __shared__ double red[2][4];
    const Geom &g = jd.g;
    const int row0 = 4 * (lane >> 4);
    const int cx = ci * 16 + (lane & 15);
    const bool diag = (bi == ci);
    const int nv = g.nvir;
    if (WRITEF) {
        const int mat[8] = {0, 2, 4, 6, 1, 3, 5, 7};
        if (w == 0 || bi != ci) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int bx = bi * 16 + row0 + r;
                if (bx < nv && cx < nv) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        fo[(w == 0) ? f_index(mat[q], bx, cx, nv, ftiled) : f_index(mat[q], cx, bx, nv, ftiled)] =
                            (double)facc[q][r];
                    }
                }
            }
        }
    }
    if (TENGY) {
        // wave 0 keeps rows r = 0,1 and receives the mirrored values for them;
        // wave 1 keeps rows r = 2,3 and receives the direct values.  Diagonal
        // pairs: wave 0's products go to an LDS tile T[q][row][col] that both
        // waves read as (b,c) and transposed as (c,b).
        T(*Tt)[16][16] = reinterpret_cast<T(*)[16][16]>(&xch[0][0][0][0]);
        if (diag) {
            if (w == 0) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) Tt[q][row0 + r][lane & 15] = facc[q][r];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int h = 0; h < 2; ++h) xch[w][q][h][lane] = facc[q][(w == 0 ? 2 : 0) + h];
        }
        if (lane < 56) *reinterpret_cast<dbl2 *>(vst + w * 112 + 2 * lane) = vreg;
        __syncthreads();
        // vst[q][s][i]: q = eorb, t1v1 t1v2 dintx1 dintx2 dintc1 dintc2
        auto vb = [&](int q, int i) { return (T)vst[q * 32 + i]; };       // b block, row i
        auto vc = [&](int q, int i) { return (T)vst[q * 32 + 16 + i]; };  // c block, column i
        const int lc = lane & 15;
        SumT<T> s = {0, 0, 0, 0};
        const T ec = vc(0, lc);
        const T u1c = vc(1, lc), u2c = vc(2, lc);
        const T x1c = vc(3, lc), x2c = vc(4, lc);
        const T y1c = vc(5, lc), y2c = vc(6, lc);
        const T eaijk = std::is_same<T, float>::value ? (T)tk.eaijk32 : (T)tk.eaijk;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = (w == 0 ? 0 : 2) + h;
            T dv[8], mv[8];
            if (diag) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    dv[q] = Tt[q][row0 + r][lane & 15];
                    mv[q] = Tt[q][lane & 15][row0 + r];
                }
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const T own = facc[q][r], other = xch[1 - w][q][h][lane];
                    dv[q] = (w == 0) ? own : other;
                    mv[q] = (w == 0) ? other : own;
                }
            }
            const int bx = bi * 16 + row0 + r;
            if (bx < nv && cx < nv) {
                const int lb = row0 + r;
                const T dd = (T)(-1.0 / (double)(vb(0, lb) + ec + eaijk));
                tengy_point<T>(dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                               mv[0], mv[1], mv[2], mv[3], mv[4], mv[5], mv[6], mv[7],
                               dd, vb(1, lb), vb(2, lb), x1c, x2c, y1c, y2c, s);
                if (bi != ci)
                    tengy_point<T>(mv[0], mv[1], mv[2], mv[3], mv[4], mv[5], mv[6], mv[7],
                                   dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                                   dd, u1c, u2c, vb(3, lb), vb(4, lb), vb(5, lb), vb(6, lb), s);
            }
        }
        s.e4i = wave_sum(s.e4i);
        s.e5i = wave_sum(s.e5i);
        s.e4k = wave_sum(s.e4k);
        s.e5k = wave_sum(s.e5k);
        if (lane == 0) {
            red[w][0] = s.e4i; red[w][1] = s.e5i; red[w][2] = s.e4k; red[w][3] = s.e5k;
        }
        __syncthreads();
        if (threadIdx.x < 4) part4[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x];
    }
    __syncthreads();  // the LDS buffers are reused by the workgroup's next item
}

// One two-wave work item: the 16x16 block pair (bi, ci) of one task.  S = 0
// takes the FP64 panels (k_fused2), 3..7 the INT8 digit planes of the split
// form (k_fused_split; 3 + kSlicesOccDigits with occupied digit planes),
// kSlicesBf16x3 the BF16 planes (k_fused_bf16x3) and kSlicesFp32 the FP32
// panels (k_fused32).
template <int S, bool TENGY, bool WRITEF>
__device__ __forceinline__ void pair_item(const JobDev &jd, const Task &tk, const PanelPairs &pp,
                                          int bi, int ci, int w, int lane,
                                          double *__restrict__ part4, double *__restrict__ fo,
                                          bool ftiled)
{
    const bool diag = (bi == ci);
    __shared__ double lds_x[2 * 8 * 2 * 64];
    __shared__ __attribute__((aligned(16))) double lds_v[7 * 2 * 16];   // energy vectors
    __shared__ __attribute__((aligned(16))) double lds_s[2][2 * 8 * 16];  // split: row scales
    const dbl2 vreg = TENGY ? stage_vectors(jd, tk, bi, ci, w, lane) : dbl2{0.0, 0.0};
    if constexpr (S == kSlicesFp32) {
        flt4 f32[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) f32[q] = flt4{0.f, 0.f, 0.f, 0.f};
        const float *pan32 = reinterpret_cast<const float *>(jd.pan);
        if (w == 0)
            fp32_wave<0>(f32, pan32, jd.g, pp, bi, ci, lane);
        else if (!diag)
            fp32_wave<1>(f32, pan32, jd.g, pp, bi, ci, lane);
        pair_epilogue<float, TENGY, WRITEF>(jd, tk, f32, bi, ci, w, lane, part4, fo, lds_x, vreg, lds_v,
                                            ftiled);
    } else {
        dbl4 facc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) facc[q] = dbl4{0.0, 0.0, 0.0, 0.0};
        if constexpr (S == 0) {
            if (w == 0)
                fp64_wave<0>(facc, jd.pan, jd.g, pp, bi, ci, lane);
            else if (!diag)
                fp64_wave<1>(facc, jd.pan, jd.g, pp, bi, ci, lane);
        } else {
            // INT8 digits (S planes) or the BF16x3 planes
            constexpr bool BF = (S == kSlicesBf16x3);
            constexpr bool T8 = (S == 3 + kSlicesOccDigits);
            constexpr int NP = BF ? 3 : T8 ? 3 : S;
            const SplitGeom &sg = jd.sg;
            const __amdgpu_buffer_rsrc_t rd = buffer_rsrc(reinterpret_cast<const double *>(jd.digits),
                                                          (uint32_t)(sg.npanels * NP * sg.slice_bytes));
            const __amdgpu_buffer_rsrc_t rt = buffer_rsrc(jd.tail, (uint32_t)(sg.npanels * sg.tail_sz * 8));
            if (w == 0)
                split_wave<NP, 0, BF, T8>(facc, rd, rt, jd.scales, jd.g, sg, pp, bi, ci, lane, lds_x, lds_s[0]);
            else if (!diag)
                split_wave<NP, 1, BF, T8>(facc, rd, rt, jd.scales, jd.g, sg, pp, bi, ci, lane, lds_x + 1024,
                                          lds_s[1]);
        }
        pair_epilogue<double, TENGY, WRITEF>(jd, tk, facc, bi, ci, w, lane, part4, fo, lds_x, vreg, lds_v,
                                             ftiled);
    }
}

// Item loop of the two-wave kernels (128 threads, one block pair per item).
template <int S, bool TENGY, bool WRITEF, bool PERSIST>
__device__ __forceinline__ void pair_items(const JobDev *__restrict__ jobs,
                                           const Segment *__restrict__ segs,
                                           WorkHdr *hdr, int64_t fstride)
{
    const int nsegs = hdr->nsegs;
    const ItemRange ir = item_range(hdr->total, blockIdx.x);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // pair_item ends with a workgroup barrier, which next_item<2> relies on
    for (int item = PERSIST ? next_item<2>(hdr, ir) : plain_item(ir); item >= 0;
         item = PERSIST ? next_item<2>(hdr, ir) : -1) {
        const int si = uniform(find_segment(segs, nsegs, item));
        const Segment sg = segs[si];
        const JobDev jd = jobs[uniform(sg.job)];
        int t, p;
        decode_item(sg, jd.g.npairs, item, t, p);
        const Task tk = read_task(sg.tasks + t);
        const short2 pr = read_pair(jd.pairs + p);
        // an opaque copy of the lane index: keeps the compiler from hoisting
        // lane-dependent offsets out of the item loop, where they would hold
        // VGPRs through the whole item
        int lane_i = lane;
        if (PERSIST) asm volatile("" : "+v"(lane_i));
        pair_item<S, TENGY, WRITEF>(jd, tk, panel_pairs(tk), uniform(pr.x), uniform(pr.y), w, lane_i,
                                    sg.partial + ((int64_t)t * jd.g.npairs + p) * 4,
                                    WRITEF ? sg.fout + (int64_t)t * fstride : nullptr, sg.ftiled != 0);
    }
}

// The split form's work item (one 16x16 block pair of one task) on two waves.
// S = 3..7, or 3 + kSlicesOccDigits (S = 3 with occupied digit planes, held to
// three waves per SIMD: the compiler's own allocation lands a few VGPRs above
// the 168 that allow it).
template <int S, bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(S == 3 + kSlicesOccDigits ? 3 : 1)))
void k_fused_split(const JobDev *__restrict__ jobs, const Segment *__restrict__ segs, WorkHdr *hdr,
                   int64_t fstride)
{
    pair_items<S, TENGY, WRITEF, PERSIST>(jobs, segs, hdr, fstride);
}

// The BF16x3 split form (three BF16 planes, FP32 accumulation, FP64 recombination).
template <bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(128) void k_fused_bf16x3(const JobDev *__restrict__ jobs,
                                                      const Segment *__restrict__ segs,
                                                      WorkHdr *hdr, int64_t fstride)
{
    pair_items<kSlicesBf16x3, TENGY, WRITEF, PERSIST>(jobs, segs, hdr, fstride);
}

// The FP32 form (FP32 panels, MFMA and energy summand; FP64 partial sums).
template <bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(128) void k_fused32(const JobDev *__restrict__ jobs,
                                                 const Segment *__restrict__ segs,
                                                 WorkHdr *hdr, int64_t fstride)
{
    pair_items<kSlicesFp32, TENGY, WRITEF, PERSIST>(jobs, segs, hdr, fstride);
}

// The FP64 form's work item on two waves (the alternative tile of k_fused).
template <bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(128) void k_fused2(const JobDev *__restrict__ jobs,
                                                const Segment *__restrict__ segs,
                                                WorkHdr *hdr, int64_t fstride)
{
    pair_items<0, TENGY, WRITEF, PERSIST>(jobs, segs, hdr, fstride);
}

// ------------------------------------------------------------------ split form, row-pair tile
// Work item of the row-pair tile: the 16x16 block pairs (r0, ci) and (r0+1, ci)
// of one task -- two row blocks over one column block; the second is absent when
// r0 + 1 = nb -- on four waves, one product group each: wave 0 the n products at
// (b,c) (LJ, LK x R1, R3), wave 1 the t products (LJt, LKt x R1t, R3t), waves 2
// and 3 the mirrored ones f(c,b) (R1, R3 x LJ, LK and R1t, R3t x LJt, LKt).  Per
// 64-wide K step a wave loads the S digit planes of its two A panels at both row
// blocks and of its two B panels at ci -- 6S 16-byte loads -- for 4 products x 2
// blocks x S(S+1)/2 digit products: 18 loads per 48 MFMAs at S = 3, against 12 per
// 24 for the two-wave tile, where the vector-memory path, not the matrix cores,
// bounds the kernel.  Each wave holds its group's INT32 level accumulators for
// both blocks (4 S (4 x 2) VGPRs).  The products of both blocks then go to LDS
// and the energy of block m is evaluated by waves 2m and 2m+1 with the rows and
// lanes of the two-wave tile.  The f matrices are bitwise those of the two-wave
// tile (exact integer products, the same recombined values); the energy
// summands are evaluated in other instruction sequences, so terms agree with
// that tile's to rounding.

// Block rows of the row-pair items: rpair (r0, ci, pair index of (r0, ci), pair
// index of (r0+1, ci) or -1), read from the job's table.
struct RowPair {
    int r0, ci, p0, p1;
};

// A wave computes product A_a x B_b into slot a + 2b.  For the direct waves
// (A = L panels, B = R panels) slot s is f_(s+1) of its group; for the mirrored
// ones (A = R, B = L) slot a + 2b holds f_(b + 2a + 1) at (c,b), so matrix s
// sits in slot kMirSlot[s].
__device__ constexpr int kMirSlot[4] = {0, 2, 1, 3};

__device__ __forceinline__ RowPair read_rowpair(const int4 *p)
{
    const int *q = reinterpret_cast<const int *>(p);
    return RowPair{uniform(cread(q)), uniform(cread(q + 1)), uniform(cread(q + 2)), uniform(cread(q + 3))};
}


// The K = nvir digit products of one group over row blocks 0 and, with TWO,
// block 1 (the caller picks the instantiation once per item: a branch around the
// second block's products inside the loop made the compiler copy every
// accumulator at each join).  The first product of each level starts from zero
// (no separate clearing of the accumulators).
//
// Software pipeline at two waves per SIMD (one other wave to cover a load's
// latency): the B planes, used by both blocks, are double-buffered and loaded a
// whole K step ahead; each block's A planes have one buffer, reloaded for the
// next step as soon as that block's products of this step are issued, so they
// arrive while the other block's 24 products run.  With one block the A planes
// are double-buffered too.  No step is loaded twice.  The INT32 sums are exact,
// so the order of the products does not change a bit of the result.
template <int S, bool TWO>
__device__ __forceinline__ void rowpair_nvir(int4v (&iacc)[2][4][S], const __amdgpu_buffer_rsrc_t rd,
                                             const uint32_t (&oa)[2][2], const uint32_t (&ob)[2],
                                             uint32_t plane, uint32_t voff8, int nkv)
{
    const int4v z = {0, 0, 0, 0};
    int4v bx[2][S], by[2][S], a0[2][S], a1[2][S];
#define RP_LOAD_B(B_, kstep)                                                                     \
    _Pragma("unroll") for (int q = 0; q < S; ++q) {                                             \
        const uint32_t ko_ = (uint32_t)(kstep) * 1024 + q * plane;                               \
        B_[0][q] = buffer_load_i8x16(rd, voff8, ob[0] + ko_);                                    \
        B_[1][q] = buffer_load_i8x16(rd, voff8, ob[1] + ko_);                                    \
    }
#define RP_LOAD_A(A_, m, kstep)                                                                  \
    _Pragma("unroll") for (int q = 0; q < S; ++q) {                                             \
        const uint32_t ko_ = (uint32_t)(kstep) * 1024 + q * plane;                               \
        A_[0][q] = buffer_load_i8x16(rd, voff8, oa[m][0] + ko_);                                 \
        A_[1][q] = buffer_load_i8x16(rd, voff8, oa[m][1] + ko_);                                 \
    }
#define RP_MFMA(m, A_, B_, first)                                                                \
    _Pragma("unroll") for (int sa = 0; sa < S; ++sa) {                                          \
        _Pragma("unroll") for (int sb = 0; sb < S - sa; ++sb) {                                 \
            const int lv = sa + sb;                                                              \
            const bool z_ = (first) && sa == 0;                                                  \
            iacc[m][0][lv] = PlaneMath<false>::mfma(A_[0][sa], B_[0][sb], z_ ? z : iacc[m][0][lv]); \
            iacc[m][1][lv] = PlaneMath<false>::mfma(A_[1][sa], B_[0][sb], z_ ? z : iacc[m][1][lv]); \
            iacc[m][2][lv] = PlaneMath<false>::mfma(A_[0][sa], B_[1][sb], z_ ? z : iacc[m][2][lv]); \
            iacc[m][3][lv] = PlaneMath<false>::mfma(A_[1][sa], B_[1][sb], z_ ? z : iacc[m][3][lv]); \
        }                                                                                        \
    }
    if constexpr (S > 3 && TWO) {
        // four planes: a second B buffer would spill (255 VGPRs).  One buffer per
        // operand, reloaded in quarter steps: a step runs block 0 x B panel 0,
        // block 0 x panel 1, block 1 x panel 0, block 1 x panel 1; block 0's A
        // planes reload after its two quarters, B panel 0 after the third, block
        // 1's A planes and B panel 1 after the fourth, so every reload has at
        // least one quarter (20 MFMAs) to arrive before its first use
#define RP_LOAD_BP(p, kstep)                                                                     \
        _Pragma("unroll") for (int q = 0; q < S; ++q)                                           \
            bx[p][q] = buffer_load_i8x16(rd, voff8, ob[p] + (uint32_t)(kstep) * 1024 + q * plane);
#define RP_MFMA_Q(m, A_, p, first)                                                               \
        _Pragma("unroll") for (int sa = 0; sa < S; ++sa) {                                      \
            _Pragma("unroll") for (int sb = 0; sb < S - sa; ++sb) {                             \
                const int lv = sa + sb;                                                          \
                const bool z_ = (first) && sa == 0;                                              \
                iacc[m][2 * p][lv] = PlaneMath<false>::mfma(A_[0][sa], bx[p][sb], z_ ? z : iacc[m][2 * p][lv]); \
                iacc[m][2 * p + 1][lv] =                                                         \
                    PlaneMath<false>::mfma(A_[1][sa], bx[p][sb], z_ ? z : iacc[m][2 * p + 1][lv]); \
            }                                                                                    \
        }
        RP_LOAD_B(bx, 0)
        RP_LOAD_A(a0, 0, 0)
        RP_LOAD_A(a1, 1, 0)
        const bool more0 = nkv > 1;
        RP_MFMA_Q(0, a0, 0, true)
        RP_MFMA_Q(0, a0, 1, true)
        if (more0) RP_LOAD_A(a0, 0, 1)
        RP_MFMA_Q(1, a1, 0, true)
        if (more0) RP_LOAD_BP(0, 1)
        RP_MFMA_Q(1, a1, 1, true)
        if (more0) {
            RP_LOAD_A(a1, 1, 1)
            RP_LOAD_BP(1, 1)
        }
        for (int ks = 1; ks < nkv; ++ks) {
            const bool more = ks + 1 < nkv;
            RP_MFMA_Q(0, a0, 0, false)
            RP_MFMA_Q(0, a0, 1, false)
            if (more) RP_LOAD_A(a0, 0, ks + 1)
            RP_MFMA_Q(1, a1, 0, false)
            if (more) RP_LOAD_BP(0, ks + 1)
            RP_MFMA_Q(1, a1, 1, false)
            if (more) {
                RP_LOAD_A(a1, 1, ks + 1)
                RP_LOAD_BP(1, ks + 1)
            }
        }
#undef RP_LOAD_BP
#undef RP_MFMA_Q
    } else if constexpr (S > 3) {
        // four planes, one block: each step's planes loaded at its start
        RP_LOAD_B(bx, 0)
        RP_LOAD_A(a0, 0, 0)
        RP_MFMA(0, a0, bx, true)
        for (int ks = 1; ks < nkv; ++ks) {
            RP_LOAD_B(bx, ks)
            RP_LOAD_A(a0, 0, ks)
            RP_MFMA(0, a0, bx, false)
        }
    } else if constexpr (TWO) {
        RP_LOAD_B(bx, 0)
        RP_LOAD_A(a0, 0, 0)
        RP_LOAD_A(a1, 1, 0)
        if (nkv > 1) RP_LOAD_B(by, 1)
        RP_MFMA(0, a0, bx, true)
        if (nkv > 1) RP_LOAD_A(a0, 0, 1)
        RP_MFMA(1, a1, bx, true)
        if (nkv > 1) RP_LOAD_A(a1, 1, 1)
        int ks = 1;
        for (; ks + 1 < nkv; ks += 2) {
            RP_LOAD_B(bx, ks + 1)
            RP_MFMA(0, a0, by, false)
            RP_LOAD_A(a0, 0, ks + 1)
            RP_MFMA(1, a1, by, false)
            RP_LOAD_A(a1, 1, ks + 1)
            const bool more = ks + 2 < nkv;
            if (more) RP_LOAD_B(by, ks + 2)
            RP_MFMA(0, a0, bx, false)
            if (more) RP_LOAD_A(a0, 0, ks + 2)
            RP_MFMA(1, a1, bx, false)
            if (more) RP_LOAD_A(a1, 1, ks + 2)
        }
        if (ks < nkv) {
            RP_MFMA(0, a0, by, false)
            RP_MFMA(1, a1, by, false)
        }
    } else {
        // one block: ping-pong of both operands (a0/bx and a1/by), two steps per trip
        RP_LOAD_B(bx, 0)
        RP_LOAD_A(a0, 0, 0)
        if (nkv > 1) {
            RP_LOAD_B(by, 1)
            RP_LOAD_A(a1, 0, 1)
        }
        RP_MFMA(0, a0, bx, true)
        int ks = 1;
        for (; ks + 1 < nkv; ks += 2) {
            RP_LOAD_B(bx, ks + 1)
            RP_LOAD_A(a0, 0, ks + 1)
            RP_MFMA(0, a1, by, false)
            if (ks + 2 < nkv) {
                RP_LOAD_B(by, ks + 2)
                RP_LOAD_A(a1, 0, ks + 2)
            }
            RP_MFMA(0, a0, bx, false)
        }
        if (ks < nkv) RP_MFMA(0, a1, by, false)
    }
#undef RP_LOAD_B
#undef RP_LOAD_A
#undef RP_MFMA
}

// One row-pair item on four waves (see above).  part0/part1: the partials of
// block pairs (r0, ci) and (r0+1, ci) (part1 unused without the second block).
template <int S, bool TENGY, bool WRITEF, bool REGS>
__device__ __forceinline__ void rowpair_item(const JobDev &jd, const Task &tk, const PanelPairs &pp,
                                             const RowPair &rp, int w, int lane,
                                             double *__restrict__ part0, double *__restrict__ part1,
                                             double *__restrict__ fo, bool ftiled)
{
    constexpr bool T8 = (S == 3 + kSlicesOccDigits);
    constexpr int NP = T8 ? 3 : S;
    static_assert(NP <= kRowPairMaxSlices, "row-pair tile: at most 4 slices");
    // LDS: every wave's FP64 products X[w][m][slot][row][lane] (64 KiB), the row
    // scales of each wave (two sets of 96), the energy vectors of the three
    // blocks [q][block r0, r0+1, ci][16], and the per-wave energy sums
    __shared__ double X[4][2][4][4][64];
    __shared__ __attribute__((aligned(16))) double lds_s[4][2 * 96];
    __shared__ __attribute__((aligned(16))) double vst[7 * 3 * 16];
    __shared__ double red[4][4];
    const bool two = rp.p1 >= 0;
    const int r0 = rp.r0, ci = rp.ci, r1 = two ? rp.r0 + 1 : rp.r0;
    const Geom &g = jd.g;
    const SplitGeom &sg = jd.sg;
    const int nvp = g.nvp;
    // wave roles: w & 1 = group (n, t), w >> 1 = mirrored
    const int LJ = 2 * pp.li, LK = LJ + 1, LJT = 2 * pp.lk, LKT = LJT + 1;
    const int R1 = 2 * pp.rkj, R3 = R1 + 1, R1T = 2 * pp.rij, R3T = R1T + 1;
    const int gi = w & 1, mir = w >> 1;
    const int L0 = gi ? LJT : LJ, L1 = gi ? LKT : LK, Q0 = gi ? R1T : R1, Q1 = gi ? R3T : R3;
    const int A0 = mir ? Q0 : L0, A1 = mir ? Q1 : L1, B0 = mir ? L0 : Q0, B1 = mir ? L1 : Q1;
    // The item's small inputs are fetched into registers first and written to
    // LDS only after the digit loop, so their latency overlaps the loop's first
    // loads instead of preceding them behind a workgroup barrier (the energy
    // epilogue's barrier orders the vector table; each wave reads only its own
    // scales).
    // row scales: lane l < 48 fetches elements 2l, 2l+1 of the 96: A side [a][m][16],
    // then B side [b][16]
    dbl2 sreg = {0.0, 0.0}, sreg_t = {0.0, 0.0}, vreg = {0.0, 0.0};
    if (lane < 48) {
        const int e = 2 * lane;
        int pid, row;
        if (e < 64) {
            pid = (e >> 5) ? A1 : A0;
            row = (((e >> 4) & 1) ? r1 : r0) * 16 + (e & 15);
        } else {
            pid = ((e - 64) >> 4) ? B1 : B0;
            row = ci * 16 + (e & 15);
        }
        const int64_t so = (int64_t)pid * nvp + row;
        sreg = gload2(jd.scales + so);
        if (T8) sreg_t = gload2(jd.scales + (int64_t)sg.npanels * nvp + so);
    }
    // energy vectors: 7 x 3 blocks x 16 = 336 doubles, two per thread of waves 0..2
    if (TENGY && threadIdx.x < 168) {
        const int e = 2 * threadIdx.x;
        const int q = e / 48, blk = (e / 16) % 3, i = e & 15;
        int row = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r) row = q == r + 1 ? uniform(tk.vrow[r]) : row;
        const double *src = q == 0 ? jd.ev : jd.vec6 + (int64_t)row * nvp;
        const int xb = blk == 0 ? r0 : blk == 1 ? r1 : ci;
        vreg = gload2(src + xb * 16 + i);
    }
    const __amdgpu_buffer_rsrc_t rd = buffer_rsrc(reinterpret_cast<const double *>(jd.digits),
                                                  (uint32_t)(sg.npanels * NP * sg.slice_bytes));
    const uint32_t plane = (uint32_t)sg.slice_bytes;
    const uint32_t blk8 = (uint32_t)(sg.nkv + sg.nkx) * 1024;
    const uint32_t voff8 = lane * 16;
    uint32_t oa[2][2], ob[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        oa[m][0] = (uint32_t)A0 * NP * plane + (m ? r1 : r0) * blk8;
        oa[m][1] = (uint32_t)A1 * NP * plane + (m ? r1 : r0) * blk8;
    }
    ob[0] = (uint32_t)B0 * NP * plane + ci * blk8;
    ob[1] = (uint32_t)B1 * NP * plane + ci * blk8;
    int4v iacc[2][4][NP];
    if (two) rowpair_nvir<NP, true>(iacc, rd, oa, ob, plane, voff8, sg.nkv);
    else rowpair_nvir<NP, false>(iacc, rd, oa, ob, plane, voff8, sg.nkv);
    if (lane < 48) {
        *reinterpret_cast<dbl2 *>(&lds_s[w][2 * lane]) = sreg;
        if (T8) *reinterpret_cast<dbl2 *>(&lds_s[w][96 + 2 * lane]) = sreg_t;
    }
    if (TENGY && threadIdx.x < 168) *reinterpret_cast<dbl2 *>(&vst[2 * threadIdx.x]) = vreg;
    // LDS operations of one wave complete in order; the fence keeps the compiler
    // from moving the scale reads below above these writes
    asm volatile("" ::: "memory");
    // folded occupied digit planes (k_split_pack, nocc <= 32): A [d1|d2], [d3|0];
    // B [d1|0], [d2|d1], [d3|d2].  Fetched here (after the staged values
    // above, whose stores would otherwise wait for them), so they arrive while
    // the K = nvir products are recombined and stashed.
    int4v obx[2][3], oax[2][2][2];
    if constexpr (T8) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) obx[p][q] = int4v{0, 0, 0, 0};
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int p = 0; p < 2; ++p) oax[m][p][0] = oax[m][p][1] = int4v{0, 0, 0, 0};
    }
    if (T8 && sg.fold) {
        const uint32_t k0 = (uint32_t)sg.nkv * 1024, k1 = k0 + 1024;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            obx[p][0] = buffer_load_i8x16(rd, voff8, ob[p] + 2 * plane + k0);
            obx[p][1] = buffer_load_i8x16(rd, voff8, ob[p] + k1);
            obx[p][2] = buffer_load_i8x16(rd, voff8, ob[p] + plane + k1);
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                oax[m][p][0] = buffer_load_i8x16(rd, voff8, oa[m][p] + k0);
                oax[m][p][1] = buffer_load_i8x16(rd, voff8, oa[m][p] + plane + k0);
            }
    }

    const double *sst = lds_s[w];
    double *xw = &X[w][0][0][0][0];  // this wave's [m][slot][row][lane]
    const int row0 = 4 * (lane >> 4);
    const int nv = g.nvir;
    const int cx = ci * 16 + (lane & 15);
    // 2^(-7(S+1)): weight of the lowest level, applied with the column scale
    const double wlow = __builtin_ldexp(1.0, -7 * (NP + 1));
    // products f = V sA(b) (sB(c) wlow) of both blocks into fw (add: onto fw)
    dbl4 fw[2][4];
    auto recombine = [&](const int4v (&acc)[2][4][NP], const double *ss, bool add) {
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int ca = c & 1, cb = c >> 1;
                const double sB = ss[64 + cb * 16 + (lane & 15)] * wlow;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double f = level_sum<NP>(acc[m][c], r) * ss[(ca * 2 + m) * 16 + row0 + r] * sB;
                    fw[m][c][r] = add ? fw[m][c][r] + f : f;
                }
            }
    };
    recombine(iacc, sst, false);
    if constexpr (T8) {
        // the K = nocc half as digit planes with scales of their own (sst + 96).
        // Where registers do not allow the K = nvir products to stay in fw over
        // its products, they wait in this wave's LDS slots (xw is not
        // restrict-qualified, and the fence keeps the compiler from forwarding the
        // stores to the reloads, which would hold them in registers after all)
        auto stash = [&]() {
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) xw[((m * 4 + c) * 4 + r) * 64 + lane] = fw[m][c][r];
            asm volatile("" ::: "memory");
        };
        auto unstash = [&]() {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) fw[m][c][r] = xw[((m * 4 + c) * 4 + r) * 64 + lane];
        };
        const int4v z = {0, 0, 0, 0};
        if (sg.fold) {
            // the folded planes fetched after the K = nvir loop; levels as
            // split_wave's.  REGS: the K = nvir products stay in fw meanwhile
            // (236 VGPRs; the persistent kernel would spill, so it stashes)
            if constexpr (!REGS) stash();
            const int4v (&bx)[2][3] = obx;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int4v (&ax)[2][2] = oax[m];
#pragma unroll
                for (int lv = 0; lv < 3; ++lv) {
                    iacc[m][0][lv] = PlaneMath<false>::mfma(ax[0][0], bx[0][lv], z);
                    iacc[m][1][lv] = PlaneMath<false>::mfma(ax[1][0], bx[0][lv], z);
                    iacc[m][2][lv] = PlaneMath<false>::mfma(ax[0][0], bx[1][lv], z);
                    iacc[m][3][lv] = PlaneMath<false>::mfma(ax[1][0], bx[1][lv], z);
                }
                iacc[m][0][2] = PlaneMath<false>::mfma(ax[0][1], bx[0][0], iacc[m][0][2]);
                iacc[m][1][2] = PlaneMath<false>::mfma(ax[1][1], bx[0][0], iacc[m][1][2]);
                iacc[m][2][2] = PlaneMath<false>::mfma(ax[0][1], bx[1][0], iacc[m][2][2]);
                iacc[m][3][2] = PlaneMath<false>::mfma(ax[1][1], bx[1][0], iacc[m][3][2]);
            }
            if constexpr (!REGS) unstash();
        } else {
            // one digit per 64-wide K step (nocc > 32): B planes, then each block's
            // A panels in turn, the K = nvir products stashed
            stash();
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int q = 0; q < NP; ++q) iacc[m][c][q] = z;
            for (int kx = 0; kx < sg.nkx; ++kx) {
                const uint32_t ko = (uint32_t)(sg.nkv + kx) * 1024;
                int4v bq[2][NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    bq[0][q] = buffer_load_i8x16(rd, voff8, ob[0] + ko + q * plane);
                    bq[1][q] = buffer_load_i8x16(rd, voff8, ob[1] + ko + q * plane);
                }
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        int4v ax[NP];
#pragma unroll
                        for (int q = 0; q < NP; ++q) ax[q] = buffer_load_i8x16(rd, voff8, oa[m][p] + ko + q * plane);
#pragma unroll
                        for (int sa = 0; sa < NP; ++sa)
#pragma unroll
                            for (int sb = 0; sb < NP - sa; ++sb) {
                                const int lv = sa + sb;
                                iacc[m][p][lv] = PlaneMath<false>::mfma(ax[sa], bq[0][sb], iacc[m][p][lv]);
                                iacc[m][p + 2][lv] = PlaneMath<false>::mfma(ax[sa], bq[1][sb], iacc[m][p + 2][lv]);
                            }
                    }
            }
            unstash();
        }
        recombine(iacc, sst + 96, true);
    } else {
        // FP64 tail (K = nocc) in the step-pair layout, A rows fed in the order
        // 4*(i%4) + i/4 (as split_wave), accumulated onto the recombined products
        const __amdgpu_buffer_rsrc_t rt = buffer_rsrc(jd.tail, (uint32_t)(sg.npanels * sg.tail_sz * 8));
        const uint32_t tsb = (uint32_t)(sg.tail_sz * 8);
        const uint32_t blkt = (uint32_t)sg.nkt * 512;
        const int li = lane & 15;
        const uint32_t voffA = (uint32_t)((4 * (li & 3) + (li >> 2)) + (lane & 48)) * 16;
        const uint32_t voffB = lane * 16;
        for (int sp = 0; sp < sg.nkt / 2; ++sp) {
            const dbl2 b2[2] = {buffer_load_f64x2(rt, voffB, (uint32_t)B0 * tsb + ci * blkt + sp * 1024),
                                buffer_load_f64x2(rt, voffB, (uint32_t)B1 * tsb + ci * blkt + sp * 1024)};
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int bm = m ? r1 : r0;
                const dbl2 a2[2] = {buffer_load_f64x2(rt, voffA, (uint32_t)A0 * tsb + bm * blkt + sp * 1024),
                                    buffer_load_f64x2(rt, voffA, (uint32_t)A1 * tsb + bm * blkt + sp * 1024)};
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        fw[m][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[c & 1][h], b2[c >> 1][h], fw[m][c], 0, 0, 0);
            }
        }
    }
    if (WRITEF) {
        // body order f1n,f1t,f2n,f2t,f3n,f3t,f4n,f4t: slot s of group gi is matrix
        // 2s + gi (mirrored: 2 kMirSlot[s] + gi at (c,b)); on a diagonal block the
        // direct waves write every element
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int bm = m ? r1 : r0;
            if ((m == 1 && !two) || (mir && bm == ci)) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int bx = bm * 16 + row0 + r;
                if (bx < nv && cx < nv) {

#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int sm = mir ? kMirSlot[s] : s;
                        fo[mir ? f_index(2 * sm + gi, cx, bx, nv, ftiled) : f_index(2 * sm + gi, bx, cx, nv, ftiled)] =
                            fw[m][s][r];
                    }
                }
            }
        }
    }
    if (TENGY) {
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[w][m][s][r][lane] = fw[m][s][r];
        __syncthreads();
        // block pair m = w >> 1 on waves 2m, 2m+1, rows r = 2 (w & 1) + h of each
        // lane's four
        const int m = w >> 1, wo = w & 1;
        if (m == 0 || two) {
            const int bm = m ? r1 : r0;
            const bool diag = (bm == ci);
            auto vb = [&](int q, int i) { return vst[q * 48 + m * 16 + i]; };
            auto vc = [&](int q, int i) { return vst[q * 48 + 32 + i]; };
            const int lc = lane & 15;
            Sum4 s4 = {0, 0, 0, 0};
            const double ec = vc(0, lc);
            const double u1c = vc(1, lc), u2c = vc(2, lc);
            const double x1c = vc(3, lc), x2c = vc(4, lc);
            const double y1c = vc(5, lc), y2c = vc(6, lc);
            const double eaijk = tk.eaijk;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int r = 2 * wo + h;
                // dv: n1..n4 (wave 0), t1..t4 (wave 1) at (b,c); mv: the same at (c,b)
                double dv[8], mv[8];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    dv[s] = X[0][m][s][r][lane];
                    dv[4 + s] = X[1][m][s][r][lane];
                    mv[s] = X[2][m][kMirSlot[s]][r][lane];
                    mv[4 + s] = X[3][m][kMirSlot[s]][r][lane];
                }
                const int bx = bm * 16 + row0 + r;
                if (bx < nv && cx < nv) {
                    const int lb = row0 + r;
                    const double dd = -1.0 / (vb(0, lb) + ec + eaijk);
                    tengy_point<double>(dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                                        mv[0], mv[1], mv[2], mv[3], mv[4], mv[5], mv[6], mv[7],
                                        dd, vb(1, lb), vb(2, lb), x1c, x2c, y1c, y2c, s4);
                    if (!diag)
                        tengy_point<double>(mv[0], mv[1], mv[2], mv[3], mv[4], mv[5], mv[6], mv[7],
                                            dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                                            dd, u1c, u2c, vb(3, lb), vb(4, lb), vb(5, lb), vb(6, lb), s4);
                }
            }
            s4.e4i = wave_sum(s4.e4i);
            s4.e5i = wave_sum(s4.e5i);
            s4.e4k = wave_sum(s4.e4k);
            s4.e5k = wave_sum(s4.e5k);
            if (lane == 0) {
                red[w][0] = s4.e4i; red[w][1] = s4.e5i; red[w][2] = s4.e4k; red[w][3] = s4.e5k;
            }
        }
        __syncthreads();
        if (threadIdx.x < 4) part0[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x];
        else if (two && threadIdx.x < 8) part1[threadIdx.x - 4] = red[2][threadIdx.x - 4] + red[3][threadIdx.x - 4];
    }
    __syncthreads();  // the LDS buffers are reused by the workgroup's next item
}

template <int S, bool TENGY, bool WRITEF, bool PERSIST>
__device__ __forceinline__ void rowpair_items(const JobDev *__restrict__ jobs,
                                              const Segment *__restrict__ segs, WorkHdr *hdr,
                                              int64_t fstride)
{
    const int nsegs = hdr->nsegs;
    const ItemRange ir = item_range(hdr->total, blockIdx.x);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // rowpair_item ends with a workgroup barrier, which next_item relies on
    for (int item = PERSIST ? next_item<4>(hdr, ir) : plain_item(ir); item >= 0;
         item = PERSIST ? next_item<4>(hdr, ir) : -1) {
        const int si = uniform(find_segment(segs, nsegs, item));
        const Segment sg = segs[si];
        const JobDev jd = jobs[uniform(sg.job)];
        int t, p;
        decode_item(sg, jd.g.nrpairs, item, t, p);
        const Task tk = read_task(sg.tasks + t);
        const RowPair rp = read_rowpair(jd.rpairs + p);
        int lane_i = lane;
        if (PERSIST) asm volatile("" : "+v"(lane_i));
        double *pbase = sg.partial + (int64_t)t * jd.g.npairs * 4;
        rowpair_item<S, TENGY, WRITEF, !PERSIST>(jd, tk, panel_pairs(tk), rp, w, lane_i, pbase + (int64_t)rp.p0 * 4,
                                       pbase + (int64_t)(rp.p1 >= 0 ? rp.p1 : rp.p0) * 4,
                                       WRITEF ? sg.fout + (int64_t)t * fstride : nullptr, sg.ftiled != 0);
    }
}

// The Ozaki INT8 split form on the row-pair tile: S = 3 or 4, or 3 +
// kSlicesOccDigits; two waves per SIMD (212 VGPRs at S = 3, 236 at S = 4 and at
// S = 3 with occupied digit planes; the 74.5 KB of LDS also fits two workgroups
// per CU).  Forced to three waves per
// SIMD (168 VGPRs) the first version of this tile ran 11 % slower.
template <int S, bool TENGY, bool WRITEF, bool PERSIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_split_rows(const JobDev *__restrict__ jobs, const Segment *__restrict__ segs, WorkHdr *hdr,
                  int64_t fstride)
{
    rowpair_items<S, TENGY, WRITEF, PERSIST>(jobs, segs, hdr, fstride);
}

template <int S>
static hipError_t launch_split_rows_s(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                                      int grid, bool persist, int64_t fstride, bool tengy,
                                      bool writef, hipStream_t st)
{
    dim3 g3((unsigned)grid), block(256);
    if (persist && tengy && !writef)
        hipLaunchKernelGGL((k_split_rows<S, true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (persist)
        return hipErrorInvalidValue;
    else if (tengy && writef)
        hipLaunchKernelGGL((k_split_rows<S, true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (tengy)
        hipLaunchKernelGGL((k_split_rows<S, true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (writef)
        hipLaunchKernelGGL((k_split_rows<S, false, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

static hipError_t launch_fused2(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                                int grid, bool persist, int64_t fstride, bool tengy, bool writef,
                                hipStream_t st)
{
    dim3 g3((unsigned)grid), block(128);
    if (persist && tengy && !writef)
        hipLaunchKernelGGL((k_fused2<true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (persist)
        return hipErrorInvalidValue;
    else if (tengy && writef)
        hipLaunchKernelGGL((k_fused2<true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (tengy)
        hipLaunchKernelGGL((k_fused2<true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (writef)
        hipLaunchKernelGGL((k_fused2<false, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int S>
static hipError_t launch_split_s(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                                 int grid, bool persist, int64_t fstride, bool tengy, bool writef,
                                 hipStream_t st)
{
    dim3 g3((unsigned)grid), block(128);
    if (persist && tengy && !writef)
        hipLaunchKernelGGL((k_fused_split<S, true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (persist)
        return hipErrorInvalidValue;
    else if (tengy && writef)
        hipLaunchKernelGGL((k_fused_split<S, true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (tengy)
        hipLaunchKernelGGL((k_fused_split<S, true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else if (writef)
        hipLaunchKernelGGL((k_fused_split<S, false, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fused_split(const JobDev *jobs, const Segment *segs, WorkHdr *hdr,
                              int grid, bool persist, int slices, int variant, int64_t fstride,
                              bool tengy, bool writef, hipStream_t st)
{
    if (grid <= 0) return hipErrorInvalidValue;
    if (variant == kTileRowPair) {
        switch (slices) {
        case 3 + kSlicesOccDigits:
            return launch_split_rows_s<3 + kSlicesOccDigits>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
        case 3: return launch_split_rows_s<3>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
        case 4: return launch_split_rows_s<4>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
        default: return hipErrorInvalidValue;
        }
    }
    if (slices == kSlicesFp32) {
        dim3 g3((unsigned)grid), block(128);
        if (persist && tengy && !writef)
            hipLaunchKernelGGL((k_fused32<true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else if (persist)
            return hipErrorInvalidValue;
        else if (tengy && writef)
            hipLaunchKernelGGL((k_fused32<true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else if (tengy)
            hipLaunchKernelGGL((k_fused32<true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (slices == kSlicesBf16x3) {
        dim3 g3((unsigned)grid), block(128);
        if (persist && tengy && !writef)
            hipLaunchKernelGGL((k_fused_bf16x3<true, false, true>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else if (persist)
            return hipErrorInvalidValue;
        else if (tengy && writef)
            hipLaunchKernelGGL((k_fused_bf16x3<true, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else if (tengy)
            hipLaunchKernelGGL((k_fused_bf16x3<true, false, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else if (writef)
            hipLaunchKernelGGL((k_fused_bf16x3<false, true, false>), g3, block, 0, st, jobs, segs, hdr, fstride);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    switch (slices) {
    case 3 + kSlicesOccDigits:
        return launch_split_s<3 + kSlicesOccDigits>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    case 3: return launch_split_s<3>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    case 4: return launch_split_s<4>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    case 5: return launch_split_s<5>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    case 6: return launch_split_s<6>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    case 7: return launch_split_s<7>(jobs, segs, hdr, grid, persist, fstride, tengy, writef, st);
    default: return hipErrorInvalidValue;
    }
}

// Resident workgroups per CU of each fused kernel, from the occupancy API.
int fused_blocks_per_cu(bool split, int slices, int variant)
{
    int n = 0;
    hipError_t e;
    if (split && variant == kTileRowPair) {
        switch (slices) {
        case 3 + kSlicesOccDigits:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_split_rows<3 + kSlicesOccDigits, true, false, true>, 256, 0);
            break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_split_rows<3, true, false, true>, 256, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_split_rows<4, true, false, true>, 256, 0); break;
        }
    } else if (!split && variant == 1) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused2<true, false, true>, 128, 0);
    } else if (!split) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<true, false, true>, 64, 0);
    } else {
        switch (slices) {
        case kSlicesBf16x3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_bf16x3<true, false, true>, 128, 0); break;
        case kSlicesFp32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused32<true, false, true>, 128, 0); break;
        case 3 + kSlicesOccDigits:
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<3 + kSlicesOccDigits, true, false, true>, 128, 0);
            break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<3, true, false, true>, 128, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<4, true, false, true>, 128, 0); break;
        case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<5, true, false, true>, 128, 0); break;
        case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<6, true, false, true>, 128, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_split<7, true, false, true>, 128, 0); break;
        }
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

// ------------------------------------------------------------------ standalone energy pass
// Reads the eight f matrices of each task from HBM (ld = nvir, body order
// f1n,f1t,f2n,f2t,f3n,f3t,f4n,f4t) in 32x32 tile pairs (BI >= CI), one
// workgroup of 256 threads each.  The direct tile is read coalesced along b,
// the mirrored tile f(c,b) along c and transposed through LDS; every thread
// reads two adjacent rows per load (16 bytes), and the vectors of both tiles
// come through LDS with one 16-byte load per thread: on gfx950 a wave-wide load
// costs the texture-data path the same ~16 cycles at 8 or 16 bytes per lane,
// and that path, not HBM, bounded the 8-byte form.  All 32 tile loads of a
// thread are issued before the first transpose; the eight mirrored tiles are
// transposed through LDS together, behind one barrier (66 KiB, two workgroups
// per CU as the registers allow).  LOADS, chosen at launch from nvir: 0 (even):
// every row pair a thread loads lies inside the matrix, one 16-byte load each.
// 1 (odd, > 1): a pair may start at the last row b; it is loaded one row earlier
// (rows b-1, b; element b sits at a byte offset of at least 8, as b >= 1) and its
// upper element kept, again one 16-byte load.  2 (nvir = 1): single 8-byte
// loads.  (A per-lane choice between a 16- and an 8-byte load is issued as both
// loads and a select: 64 load instructions per thread, more than the wait
// counter holds, so the first LDS write would wait for all of them.)
__host__ __device__ inline int tengy_tiles(int nv) { return (nv + 31) / 32; }

int tengy_npairs(const Geom &g)
{
    const int nt = tengy_tiles(g.nvir);
    return nt * (nt + 1) / 2;
}

// DIRECT: one triple's inputs straight from the caller's device arrays (the
// allocation-free entry point, TengyDirect): eight separate f matrices, the
// virtual orbital energies and the six vectors unpadded (elements past nvir read
// as 0, as the padding of the job's own arrays does), eaijk by value.  Otherwise
// the job's f workspace, padded vector pool and task records.  Both forms run the
// same arithmetic, so their results are bitwise the same.
template <bool DIRECT, int LOADS>
__global__ __launch_bounds__(256) void k_tengy(const double *__restrict__ f, int64_t fstride, int tiled,
                                               Geom g, const Task *__restrict__ tasks,
                                               int ntasks, const double *__restrict__ ev,
                                               const double *__restrict__ vec6,
                                               double *__restrict__ partial, TengyDirect dir)
{
    __shared__ double tile[8][32][33];
    __shared__ __attribute__((aligned(16))) double vst[7 * 2 * 32];
    __shared__ double red[4][4];
    const int nv = g.nvir, nvp = g.nvp;
    const int nt = tengy_tiles(nv);
    const int np = nt * (nt + 1) / 2;
    const int L = xcd_remap(blockIdx.x, ntasks * np);
    const int t = L / np;
    int p = L - t * np;
    int BI = 0;
    while ((BI + 1) * (BI + 2) / 2 <= p) ++BI;
    const int CI = p - BI * (BI + 1) / 2;
    double eaijk;
    int vrow[6];
    if constexpr (DIRECT) {
        eaijk = dir.eaijk;
#pragma unroll
        for (int r = 0; r < 6; ++r) vrow[r] = r;
    } else {
        const Task tk = read_task(tasks + t);
        eaijk = tk.eaijk;
#pragma unroll
        for (int r = 0; r < 6; ++r) vrow[r] = uniform(tk.vrow[r]);
    }
    const double *ft = f + (int64_t)t * fstride;
    const int64_t vv = (int64_t)nv * nv;
    // thread (tx, ty): rows 2tx, 2tx+1 of the B tile, columns ty, ty+16 of the C tile
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;

    // eorb and the six vectors of both tiles, one 16-byte load per thread:
    // vst[q*64 + side*32 + i] = vector q (eorb, t1v1 t1v2 dintx1 dintx2 dintc1
    // dintc2) at row (side ? CI : BI)*32 + i; rows past nvp read as 0
    dbl2 vreg = {0.0, 0.0};
    {
        const int e = 2 * (int)threadIdx.x;
        if (e < 7 * 64) {
            const int q = e >> 6, side = (e >> 5) & 1, i = e & 31;
            const int x = (side ? CI : BI) * 32 + i;
            int row = 0;
#pragma unroll
            for (int r = 0; r < 6; ++r) row = q == r + 1 ? vrow[r] : row;
            if constexpr (DIRECT) {
                // unpadded caller arrays (x is even): both, one or neither element
                const double *src = dir.v[0];
#pragma unroll
                for (int r = 1; r < 7; ++r) src = q == r ? dir.v[r] : src;
                if (x + 1 < nv) vreg = gload2(src + x);
                else if (x < nv) vreg = dbl2{gld(src + x), 0.0};
            } else {
                const double *src = q == 0 ? ev : vec6 + (int64_t)row * nvp;
                if (x < nvp) vreg = gload2(src + x);  // nvp is a multiple of 16: x+1 < nvp too
            }
        }
    }

    // Direct values d[m][q] = f_m(bx, bx+1; cx), bx = BI*32 + 2tx, cx = CI*32 + ty + 16q;
    // mirrored m[m][q] = f_m(cc, cc+1; bb), cc = CI*32 + 2tx, bb = BI*32 + ty + 16q,
    // transposed below.  Byte offsets inside one matrix; the matrix is selected by
    // the uniform offset m*vv*8.  Out-of-range points load from offset 0 and are
    // masked; a pair whose second row is past nv (odd nv) is read one row early.
    // one resource over the eight matrices (matrix m at byte offset m vv 8), or
    // one per matrix for DIRECT (each at offset 0)
    __amdgpu_buffer_rsrc_t rf = buffer_rsrc(ft, (uint32_t)(8 * vv * 8));
    __amdgpu_buffer_rsrc_t rfm[8];
    if constexpr (DIRECT) {
#pragma unroll
        for (int m = 0; m < 8; ++m) rfm[m] = buffer_rsrc(dir.f[m], (uint32_t)(vv * 8));
    }
    uint32_t od[2], om[2];
    bool okd[2], okm[2];
    const int r0 = 2 * tx;
    const bool two_d = BI * 32 + r0 + 1 < nv, two_m = CI * 32 + r0 + 1 < nv;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int bx = BI * 32 + r0, cx = CI * 32 + ty + 16 * q;
        okd[q] = bx < nv && cx < nv;
        od[q] = okd[q] ? (uint32_t)f_index(0, bx, cx, nv, !DIRECT && tiled) * 8 : 0;
        const int cc = CI * 32 + r0, bb = BI * 32 + ty + 16 * q;
        okm[q] = cc < nv && bb < nv;
        om[q] = okm[q] ? (uint32_t)f_index(0, cc, bb, nv, !DIRECT && tiled) * 8 : 0;
    }
    // bytes from one matrix to the next: nv^2 doubles column-major, one tile in
    // the tiled layout (f_index)
    const bool tl = !DIRECT && tiled;
    const uint32_t mstride = tl ? (uint32_t)((nv - 32 * BI < 32 ? nv - 32 * BI : 32) *
                                             (nv - 32 * CI < 32 ? nv - 32 * CI : 32) * 8)
                                : (uint32_t)(vv * 8);
    auto load_pair = [&](int m, uint32_t off, bool two) -> dbl2 {
        const __amdgpu_buffer_rsrc_t r = DIRECT ? rfm[m] : rf;
        const uint32_t moff = DIRECT ? 0u : (uint32_t)m * mstride;
        if constexpr (LOADS == 0) {
            return buffer_load_f64x2(r, off, moff);
        } else if constexpr (LOADS == 1) {
            const dbl2 v = buffer_load_f64x2(r, two ? off : off - 8, moff);
            return two ? v : dbl2{v.y, 0.0};
        } else {
            return dbl2{buffer_load_f64(r, off, moff), 0.0};
        }
    };
    // the mirrored tiles first: their transposes can start while the direct
    // tiles are still arriving
    dbl2 dv[8][2], mv[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
            mv[m][q] = load_pair(m, om[q], okm[q] ? two_m : nv > 1);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
            dv[m][q] = load_pair(m, od[q], okd[q] ? two_d : nv > 1);
    if (2 * (int)threadIdx.x < 7 * 64) *reinterpret_cast<dbl2 *>(vst + 2 * threadIdx.x) = vreg;
    // transpose: thread (tx, ty) loaded f(c = CI*32 + 2tx + h, b = BI*32 + ty + 16q)
    // and needs f(c = CI*32 + ty + 16q, b = BI*32 + 2tx + h); all eight tiles go
    // through LDS at once, each written as its loads land, behind one barrier
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < 2; ++h) tile[m][ty + 16 * q][r0 + h] = okm[q] ? mv[m][q][h] : 0.0;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < 2; ++h) mv[m][q][h] = tile[m][r0 + h][ty + 16 * q];
    auto vb = [&](int q, int i) { return vst[q * 64 + i]; };       // B tile, row i
    auto vc = [&](int q, int i) { return vst[q * 64 + 32 + i]; };  // C tile, column i
    Sum4 s = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int lc = ty + 16 * q, cx = CI * 32 + lc;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lb = r0 + h, bx = BI * 32 + lb;
            if (bx < nv && cx < nv) {
                const double d = -1.0 / (vb(0, lb) + vc(0, lc) + eaijk);
                // body order: 0 f1n, 1 f1t, 2 f2n, 3 f2t, 4 f3n, 5 f3t, 6 f4n, 7 f4t
                tengy_point(dv[0][q][h], dv[2][q][h], dv[4][q][h], dv[6][q][h], dv[1][q][h],
                            dv[3][q][h], dv[5][q][h], dv[7][q][h], mv[0][q][h], mv[2][q][h],
                            mv[4][q][h], mv[6][q][h], mv[1][q][h], mv[3][q][h], mv[5][q][h],
                            mv[7][q][h], d, vb(1, lb), vb(2, lb), vc(3, lc), vc(4, lc), vc(5, lc),
                            vc(6, lc), s);
                if (BI != CI)
                    tengy_point(mv[0][q][h], mv[2][q][h], mv[4][q][h], mv[6][q][h], mv[1][q][h],
                                mv[3][q][h], mv[5][q][h], mv[7][q][h], dv[0][q][h], dv[2][q][h],
                                dv[4][q][h], dv[6][q][h], dv[1][q][h], dv[3][q][h], dv[5][q][h],
                                dv[7][q][h], d, vc(1, lc), vc(2, lc), vb(3, lb), vb(4, lb),
                                vb(5, lb), vb(6, lb), s);
            }
        }
    }
    s.e4i = wave_sum(s.e4i);
    s.e5i = wave_sum(s.e5i);
    s.e4k = wave_sum(s.e4k);
    s.e5k = wave_sum(s.e5k);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = s.e4i; red[w][1] = s.e5i; red[w][2] = s.e4k; red[w][3] = s.e5k;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        partial[(int64_t)L * 4 + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
    }
}

hipError_t launch_tengy(const double *f, int64_t fstride, bool tiled, const Geom &g, const Task *tasks,
                        int ntasks, const double *ev, const double *vec6, double *partial,
                        hipStream_t st)
{
    const int64_t total = (int64_t)ntasks * tengy_npairs(g);
    if (total <= 0 || total > 0x7fffffff) return hipErrorInvalidValue;
    if ((int64_t)g.nvir * g.nvir * 64 >= (int64_t(1) << 32)) return hipErrorInvalidValue;  // 32-bit offsets
    if (g.nvir % 2 == 0)
        hipLaunchKernelGGL((k_tengy<false, 0>), dim3((unsigned)total), dim3(256), 0, st, f, fstride,
                           tiled ? 1 : 0, g, tasks, ntasks, ev, vec6, partial, TengyDirect{});
    else if (g.nvir > 1)
        hipLaunchKernelGGL((k_tengy<false, 1>), dim3((unsigned)total), dim3(256), 0, st, f, fstride,
                           tiled ? 1 : 0, g, tasks, ntasks, ev, vec6, partial, TengyDirect{});
    else
        hipLaunchKernelGGL((k_tengy<false, 2>), dim3((unsigned)total), dim3(256), 0, st, f, fstride,
                           tiled ? 1 : 0, g, tasks, ntasks, ev, vec6, partial, TengyDirect{});
    return hipGetLastError();
}

// ------------------------------------------------------------------ allocation-free energy pass
// The per-tile partials of ccsdt_tengy_device_noalloc live in this file-scope
// array, one copy per device, held until the process ends.  Host code must not
// pass `scratch` itself to a kernel (in host code the name is a stand-in, not a
// device address): scratch_address() gives the current device's address, which
// the library fetches once per device and passes to the kernels.
static __device__ __align__(256) double scratch[kNoAllocPartials];

hipError_t scratch_address(double **ptr)
{
    return hipGetSymbolAddress(reinterpret_cast<void **>(ptr), HIP_SYMBOL(scratch));
}

hipError_t launch_tengy_direct(const TengyDirect &in, const Geom &g, double *scratch_ptr, double *out4,
                               hipStream_t st)
{
    const int np = tengy_npairs(g);
    if (np <= 0 || np * 4 > kNoAllocPartials) return hipErrorInvalidValue;
    if (g.nvir % 2 == 0)
        hipLaunchKernelGGL((k_tengy<true, 0>), dim3((unsigned)np), dim3(256), 0, st, nullptr, (int64_t)0, 0,
                           g, nullptr, 1, nullptr, nullptr, scratch_ptr, in);
    else if (g.nvir > 1)
        hipLaunchKernelGGL((k_tengy<true, 1>), dim3((unsigned)np), dim3(256), 0, st, nullptr, (int64_t)0, 0,
                           g, nullptr, 1, nullptr, nullptr, scratch_ptr, in);
    else
        hipLaunchKernelGGL((k_tengy<true, 2>), dim3((unsigned)np), dim3(256), 0, st, nullptr, (int64_t)0, 0,
                           g, nullptr, 1, nullptr, nullptr, scratch_ptr, in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_terms(scratch_ptr, np, 1, out4, st);
}

// ------------------------------------------------------------------ reductions
// Per task: sum of the tile partials in a fixed order (lane-strided, then
// wave_sum), so the result does not depend on scheduling.
__global__ __launch_bounds__(64) void k_reduce_terms(const double *__restrict__ partial,
                                                     int npairs, double *__restrict__ terms)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const double *pp = partial + (int64_t)t * npairs * 4;
    for (int p = lane; p < npairs; p += 64) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += pp[p * 4 + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = wave_sum(s[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) terms[(int64_t)t * 4 + q] = s[q];
    }
}

hipError_t launch_reduce_terms(const double *partial, int npairs, int ntasks, double *terms,
                               hipStream_t st)
{
    if (ntasks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reduce_terms, dim3(ntasks), dim3(64), 0, st, partial, npairs, terms);
    return hipGetLastError();
}

// Per-task sums over the segments of a work list (fixed order, as above); the
// terms go to each segment's terms array.  grid may be smaller than the task
// count (the workgroups then stride).
__global__ __launch_bounds__(64) void k_reduce_segments(const JobDev *__restrict__ jobs,
                                                        const Segment *__restrict__ segs,
                                                        const WorkHdr *__restrict__ hdr)
{
    const WorkHdr h = *hdr;
    const int lane = threadIdx.x;
    int sgi = 0, before = 0;
    for (int gt = blockIdx.x; gt < h.ntasks; gt += gridDim.x) {
        // segment of global task gt (segments hold consecutive task ranges)
        while (sgi + 1 < h.nsegs && gt >= before + segs[sgi].ntasks) {
            before += segs[sgi].ntasks;
            ++sgi;
        }
        const Segment sg = segs[sgi];
        const int npairs = jobs[sg.job].g.npairs;
        const int t = gt - before;
        const double *pp = sg.partial + (int64_t)t * npairs * 4;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int p = lane; p < npairs; p += 64) {
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] += pp[p * 4 + q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = wave_sum(s[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sg.terms[(int64_t)t * 4 + q] = s[q];
        }
    }
}

hipError_t launch_reduce_segments(const JobDev *jobs, const Segment *segs, const WorkHdr *hdr,
                                  int grid, hipStream_t st)
{
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reduce_segments, dim3(grid), dim3(64), 0, st, jobs, segs, hdr);
    return hipGetLastError();
}

// Neumaier compensated addition.
__device__ __forceinline__ void neu_add(double &sum, double &comp, double x)
{
    const double t = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
}

// Totals over the list with the accumulation rule of
// omp/ccsd_trpdrv_omp_cbody.c:133-139 (k terms only when i != k), compensated.
__global__ __launch_bounds__(256) void k_reduce_list(const double *__restrict__ terms,
                                                     const Task *__restrict__ tasks, int ntasks,
                                                     double *__restrict__ out2)
{
    __shared__ double sh[256][4];
    const int tid = threadIdx.x;
    const int chunk = (ntasks + 255) / 256;
    const int lo = tid * chunk, hi = min(ntasks, lo + chunk);
    double s4 = 0.0, c4 = 0.0, s5 = 0.0, c5 = 0.0;
    for (int t = lo; t < hi; ++t) {
        const double *tm = terms + (int64_t)t * 4;
        neu_add(s4, c4, tm[0]);
        neu_add(s5, c5, tm[1]);
        if (tasks[t].ik_differ) {
            neu_add(s4, c4, tm[2]);
            neu_add(s5, c5, tm[3]);
        }
    }
    sh[tid][0] = s4; sh[tid][1] = c4; sh[tid][2] = s5; sh[tid][3] = c5;
    __syncthreads();
    if (tid == 0) {
        double S4 = 0.0, C4 = 0.0, S5 = 0.0, C5 = 0.0;
        for (int q = 0; q < 256; ++q) {
            neu_add(S4, C4, sh[q][0]);
            neu_add(S4, C4, sh[q][1]);
            neu_add(S5, C5, sh[q][2]);
            neu_add(S5, C5, sh[q][3]);
        }
        out2[0] = S4 + C4;
        out2[1] = S5 + C5;
    }
}

hipError_t launch_reduce_list(const double *terms, const Task *tasks, int ntasks, double *out2,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_list, dim3(1), dim3(256), 0, st, terms, tasks, ntasks, out2);
    return hipGetLastError();
}

// ------------------------------------------------------------------ layout probe
// One v_mfma_f64_16x16x4_f64 on A (16x4, row major) and B (4x16, row major),
// fed as lane l: A[l%16][l/16], B[l/16][l%16]; d[4*l + r] is the lane's result,
// which is D[l/16 + 4*r][l%16].
__global__ __launch_bounds__(64) void k_mfma_probe(const double *a, const double *b, double *d)
{
    const int l = threadIdx.x;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[(l & 15) * 4 + (l >> 4)], b[(l >> 4) * 16 + (l & 15)],
                                              acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

hipError_t launch_mfma_probe(const double *a, const double *b, double *d, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, a, b, d);
    return hipGetLastError();
}

typedef int int4v __attribute__((ext_vector_type(4)));
typedef int int16v __attribute__((ext_vector_type(16)));

// One INT8 MFMA on raw per-lane operand registers: a, b hold 16 bytes per lane
// (64 lanes), d receives the lane's 4 (16x16x64) or 16 (32x32x32) results.
__global__ __launch_bounds__(64) void k_mfma_i8_probe(const int *a, const int *b, int *d, int shape)
{
    const int l = threadIdx.x;
    const int4v av = {a[4 * l], a[4 * l + 1], a[4 * l + 2], a[4 * l + 3]};
    const int4v bv = {b[4 * l], b[4 * l + 1], b[4 * l + 2], b[4 * l + 3]};
    if (shape == 0) {
        int4v acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
    } else if (shape == 3) {
        // v_mfma_f32_16x16x4_f32: the first word of each lane's operand as a float
        const flt4 acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, av[0]),
                                                              __builtin_bit_cast(float, bv[0]),
                                                              flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        float *df = reinterpret_cast<float *>(d);
#pragma unroll
        for (int r = 0; r < 4; ++r) df[l * 4 + r] = acc[r];
    } else if (shape == 2) {
        // the 16 bytes per lane as 8 BF16 values; FP32 results returned as bits
        const flt4 acc = PlaneMath<true>::mfma(av, bv, flt4{0.f, 0.f, 0.f, 0.f});
        float *df = reinterpret_cast<float *>(d);
#pragma unroll
        for (int r = 0; r < 4; ++r) df[l * 4 + r] = acc[r];
    } else {
        int16v acc = {};
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; ++r) d[l * 16 + r] = acc[r];
    }
}

hipError_t launch_mfma_i8_probe(const int *a, const int *b, int *d, int shape, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_i8_probe, dim3(1), dim3(64), 0, st, a, b, d, shape);
    return hipGetLastError();
}

// INT8 MFMA rate probe on register operands: every wave runs iters x 8 MFMAs
// (v_mfma_i32_16x16x64_i8, or with BIG v_mfma_i32_32x32x32_i8) over eight
// operand register pairs, all ones or (RND) per-lane xorshift values, so that
// consecutive MFMAs see the same or different bits.  Measures how much of the
// datasheet INT8 rate the instruction sustains for the data it is given.
__device__ __forceinline__ int4v probe_rnd4(uint32_t s)
{
    int4v v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        v[i] = (int)s;
    }
    return v;
}

template <bool BIG, bool RND>
__global__ __launch_bounds__(256) void k_mfma_rate(int *__restrict__ out, int iters)
{
    const uint32_t lane = threadIdx.x & 63;
    int4v a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        a[u] = RND ? probe_rnd4(0x9E3779B9u * (lane + 1) + 977u * u + blockIdx.x) : int4v{1, 1, 1, 1};
        b[u] = RND ? probe_rnd4(0x85EBCA6Bu * (lane + 7) + 131u * u + blockIdx.x) : int4v{1, 1, 1, 1};
    }
    int s = 0;
    if constexpr (BIG) {
        int16v acc[4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u & 3] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[u], b[u], acc[u & 3], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 4; ++u) s += acc[u][lane & 15];
    } else {
        int4v acc[8] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u], b[(u + 3) & 7], acc[u], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 8; ++u) s += acc[u][lane & 3];
    }
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

hipError_t launch_mfma_rate(int *out, int shape, bool rnd, int grid, int iters, hipStream_t st)
{
    if (shape == 1) {
        if (rnd) hipLaunchKernelGGL((k_mfma_rate<true, true>), dim3(grid), dim3(256), 0, st, out, iters);
        else hipLaunchKernelGGL((k_mfma_rate<true, false>), dim3(grid), dim3(256), 0, st, out, iters);
    } else if (shape == 0) {
        if (rnd) hipLaunchKernelGGL((k_mfma_rate<false, true>), dim3(grid), dim3(256), 0, st, out, iters);
        else hipLaunchKernelGGL((k_mfma_rate<false, false>), dim3(grid), dim3(256), 0, st, out, iters);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Plain streaming copy of n doubles (HBM counter calibration: reads and writes
// exactly 8*n bytes each).
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n2)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

hipError_t launch_copy(const double *src, double *dst, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, reinterpret_cast<const double2 *>(src),
                       reinterpret_cast<double2 *>(dst), n / 2);
    return hipGetLastError();
}

// Plain streaming read of n doubles (the read-only HBM reference): block b sums
// its own contiguous chunk of the array, every thread keeping eight independent
// 16-byte loads in flight (the fastest of the read shapes tried on MI355X:
// contiguous chunks beat grid-stride loops, and the rate levels off from about
// 8192 blocks); the block writes one double, so the writes are negligible.
__global__ __launch_bounds__(256) void k_readsum(const double2 *__restrict__ src, double *__restrict__ out, int64_t n2)
{
    constexpr int U = 8;
    const int64_t per = (n2 + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per;
    const int64_t e = b + per < n2 ? b + per : n2;
    double acc[U] = {};
    int64_t i = b + threadIdx.x;
    for (; i + (U - 1) * 256 < e; i += U * 256) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * 256];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] += v[u].x + v[u].y;
    }
    for (; i < e; i += 256) acc[0] += src[i].x + src[i].y;
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) t += acc[u];
    __shared__ double red[256];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

hipError_t launch_readsum(const double *src, double *out, int64_t n, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_readsum, dim3(blocks), dim3(256), 0, st, reinterpret_cast<const double2 *>(src), out,
                       n / 2);
    return hipGetLastError();
}

// Load-rate probe: every wave streams `iters` x 8 wave-wide loads of W dwords
// per lane from its workgroup's 16 KiB region of `buf` (64 regions, 1 MiB; the
// offsets stay below 10 KiB into the region), to VGPRs (LDS = false) or, W = 4
// only, as 16-byte global->LDS DMA into eight fixed 1 KiB slots of a 32 KiB
// array.  `stride` (512) comes at run time so no load can be hoisted.
typedef __attribute__((address_space(3))) void lds_void;
template <int W, bool LDS>
__global__ __launch_bounds__(256) void k_load_rate(const uint32_t *__restrict__ buf,
                                                   uint32_t *__restrict__ out, int iters, int stride)
{
    __shared__ __attribute__((aligned(16))) uint32_t slots[4][2048];
    const __amdgpu_buffer_rsrc_t r = buffer_rsrc(reinterpret_cast<const double *>(buf), 16384u * 64u);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t base = (blockIdx.x & 63) * 16384u;
    uint32_t acc = 0;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t off = base + lane * (LDS ? 16 : 4 * W) + (uint32_t)(((it * 8 + u) * stride) & 8191);
            if constexpr (LDS) {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void *)&slots[0][u * 256], 16, off, 0, 0, 0);
            } else if constexpr (W == 1) {
                acc += __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
            } else if constexpr (W == 2) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
                acc += v[0] + 3u * v[1];
            } else {
                const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
                acc += v[0] + 3u * v[1] + 5u * v[2] + 7u * v[3];
            }
        }
    }
    if constexpr (LDS) {
        __builtin_amdgcn_s_waitcnt(0);
        acc += slots[0][lane];
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc;
}

hipError_t launch_load_rate(const uint32_t *buf, uint32_t *out, int bytes, bool lds, int grid,
                            int iters, hipStream_t st)
{
    const dim3 g(grid), b(256);
    if (lds && bytes == 16) hipLaunchKernelGGL((k_load_rate<4, true>), g, b, 0, st, buf, out, iters, 512);
    else if (!lds && bytes == 4) hipLaunchKernelGGL((k_load_rate<1, false>), g, b, 0, st, buf, out, iters, 512);
    else if (!lds && bytes == 8) hipLaunchKernelGGL((k_load_rate<2, false>), g, b, 0, st, buf, out, iters, 512);
    else if (!lds && bytes == 16) hipLaunchKernelGGL((k_load_rate<4, false>), g, b, 0, st, buf, out, iters, 512);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace ccsdt
