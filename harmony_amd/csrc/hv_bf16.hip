// BF16 precision form of the list scan: candidate distances on the matrix cores, exact FP32
// re-rank of the survivors.
//
// For a query q probing list l (centroid c), and a vector x of the list, with residuals
// rq = q - c and rx = x - c:  |q - x|^2 = |rq|^2 + |rx|^2 - 2 rq.rx.  The residual codes rx are
// stored in BF16 (d zero-padded to dpad = 32 * NKS) with |rx|^2 in FP32; |rq|^2 is the coarse
// distance of (q, l). rq.rx is accumulated in FP32 by v_mfma_f32_16x16x32_bf16 (16 queries x
// 16 vectors x 32 dimensions per instruction).
//
// Code layout (per list, MFMA B-operand order): vector block vb (16 vectors), K-step s (32 dims),
// lane t = 16 g + c holds dims s*32 + g*8 .. +8 of vector vb*16 + c, as 8 BF16 (16 bytes):
//   element (v, k) -> (((v / 16) * NKS + k / 32) * 64 + ((k % 32) / 8) * 16 + v % 16) * 8 + k % 8
// so one K-step of one block is one contiguous 1 KB load of the wavefront.
//
// Rounding makes the estimate d~ = |rq|^2 + |rx|^2 - 2 dot~ inexact. With the rounding errors
// eq = rq - bf16(rq), ex = rx - bf16(rx) (their norms measured, plus 2^-23 |r| for the FP32
// subtraction that formed r), rq.rx - bf16(rq).bf16(rx) = eq.rx + bf16(rq).ex, so by
// Cauchy-Schwarz
//   |rq.rx - dot~| <= |eq| |rx| + (|rq| + |eq|) |ex| + g_d (|rq| + |eq|) (|rx| + |ex|)
// (g_d: FP32 accumulation of d exact BF16 products, taken twice for the MFMA's order). Every
// pair's error is bounded by
//   err = kappa1 (|eq| |rx| + (|rq| + |eq|) |ex|) + kappa_g (|rq| + |eq|) (|rx| + |ex|)
//         + kappa2 (|rq|^2 + |rx|^2)
// with kappa1 = 2 * 1.01, kappa_g = 2 * 1.01 * 2 g_d and kappa2 covering the FP32 norms, the
// FP32 evaluation of d~ and the gap between the real distance and the FP32 chain the exact path
// computes (the host derives them from d). |ex| of every vector is stored with the codes, |eq|
// of every (query, list) slot comes with the query residuals: measured errors are ~0.4 of the
// worst case 2^-8 |r| on typical data, so the bound is ~2.6x tighter than a worst-case one. So
// lb = d~ - err <= exact <= d~ + err = ub for every pair, and a pair can be dropped when
// lb > bound without changing the result. Survivors are re-ranked with the exact FP32 distance
// of the plain path, so the BF16 form returns the same ids and distances as the plain path.
#include <algorithm>

#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// Operand types of the filter scan: 8 residual values per lane and K-step, one 16x16x32 MFMA.
// BF16 codes are the residuals as they are; FP8 (e4m3) codes are scaled by a power of two per
// list (vectors) and per (query, list) slot (queries), undone on the FP32 product.
struct ElemBf16 {
    typedef __bf16 elem;
    typedef bf16x8 vec;
    static constexpr bool kScaled = false;
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
struct ElemFp8 {
    typedef uint8_t elem;
    typedef long vec;
    static constexpr bool kScaled = true;
    static __device__ __forceinline__ floatx4 mma(vec a, vec b, floatx4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
    }
};

// 8 floats -> 8 e4m3 bytes (round to nearest even, the converter's format is the MFMA's)
__device__ __forceinline__ long pack_fp8x8(const float (&v)[8]) {
    unsigned int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0u, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
    unsigned int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0u, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
    return (long)(((unsigned long long)hi << 32) | lo);
}

// Power-of-two scale that brings a largest magnitude m into [64, 128) (e4m3 holds up to 448),
// as its inverse: codes are value * 2^e, products are multiplied by 2^-e.
__device__ __forceinline__ float fp8_inv_scale(float m) {
    if (!(m > 0.f) || !(m < 3.0e38f)) return 1.f;
    const int e = min(100, max(-100, 6 - ilogbf(m)));
    return ldexpf(1.f, -e);
}

// ---------------------------------------------------------------------------------------
// Index build: BF16 residual codes, FP32 residual norms and BF16 rounding-error norms of every
// list (one workgroup per list), and the FP32 rows in list order for the re-rank.
__global__ __launch_bounds__(256) void k_build_resid(const float* __restrict__ x, int d, int dpad,
                                                     const int* __restrict__ order,
                                                     const int64_t* __restrict__ list_off,
                                                     const float* __restrict__ centroids,
                                                     const int64_t* __restrict__ roff, __bf16* __restrict__ rcodes,
                                                     float* __restrict__ rnorm, float* __restrict__ rerr) {
    const int l = blockIdx.x;
    const int64_t b = list_off[l];
    const int n = (int)(list_off[l + 1] - b);
    const int n16 = (n + 15) / 16 * 16;
    const float* __restrict__ c = centroids + (int64_t)l * d;
    __bf16* __restrict__ out = rcodes + roff[l];
    const int nks = dpad / 32;
    for (int64_t e = threadIdx.x; e < (int64_t)n16 * dpad; e += blockDim.x) {
        const int v = (int)(e / dpad), k = (int)(e - (int64_t)v * dpad);
        float r = 0.f;
        if (v < n && k < d) r = x[(int64_t)order[b + v] * d + k] - c[k];
        const int64_t o = ((((int64_t)(v >> 4) * nks + (k >> 5)) * 64 + ((k & 31) >> 3) * 16 + (v & 15)) << 3) + (k & 7);
        out[o] = (__bf16)r;
    }
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
        const float* __restrict__ xv = x + (int64_t)order[b + v] * d;
        float acc = 0.f, ae = 0.f;
        for (int k = 0; k < d; ++k) {
            const float t = xv[k] - c[k];
            acc = __builtin_fmaf(t, t, acc);
            const float e = t - (float)(__bf16)t;  // exact: t and its BF16 rounding share the exponent
            ae = __builtin_fmaf(e, e, ae);
        }
        rnorm[b + v] = acc;
        rerr[b + v] = __builtin_sqrtf(ae) + 0x1p-23f * __builtin_sqrtf(acc);
    }
}

void launch_build_resid(const float* x, int d, int dpad, const int* order, const int64_t* list_off,
                        const float* centroids, const int64_t* roff, int nlist, __bf16* rcodes, float* rnorm,
                        float* rerr, hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_build_resid, dim3((unsigned)nlist), dim3(256), 0, s, x, d, dpad, order, list_off, centroids,
                       roff, rcodes, rnorm, rerr);
}

__global__ void k_gather_rows(const float* __restrict__ rows, int64_t n, int d, const int* __restrict__ order,
                              float* __restrict__ by_pos) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * d) return;
    const int64_t p = hv_idiv(e, d);
    const int k = (int)(e - p * d);
    by_pos[e] = rows[(int64_t)order[p] * d + k];
}

void launch_gather_rows(const float* rows, int64_t n, int d, const int* order, float* by_pos, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((n * d + 255) / 256)), dim3(256), 0, s, rows, n, d, order,
                       by_pos);
}

// v[e] = q[k0 + e] - c[k0 + e] (0 past d); 16-byte loads when the 8 values are in range and
// d % 4 == 0 (rows then start on 16-byte boundaries)
__device__ __forceinline__ void load8_diff(const float* __restrict__ qv, const float* __restrict__ cv, int k0, int d,
                                           float (&v)[8]) {
    if ((d & 3) == 0 && k0 + 8 <= d) {
        const float4 q0 = *reinterpret_cast<const float4*>(qv + k0);
        const float4 q1 = *reinterpret_cast<const float4*>(qv + k0 + 4);
        const float4 c0 = *reinterpret_cast<const float4*>(cv + k0);
        const float4 c1 = *reinterpret_cast<const float4*>(cv + k0 + 4);
        v[0] = q0.x - c0.x;
        v[1] = q0.y - c0.y;
        v[2] = q0.z - c0.z;
        v[3] = q0.w - c0.w;
        v[4] = q1.x - c1.x;
        v[5] = q1.y - c1.y;
        v[6] = q1.z - c1.z;
        v[7] = q1.w - c1.w;
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + e;
            v[e] = k < d ? qv[k] - cv[k] : 0.f;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Query residuals of every (query, probe rank) slot: BF16 (q - c) rows of dpad dims (zero past
// d) and the norm of their rounding error (plus 2^-23 |q - c|, see the bound above). A group of
// G lanes per slot (G = the slot's 8-dim words rounded up to a power of two, <= 64; lane g holds
// dims 8g .. 8g+7).
template <int G>
__global__ __launch_bounds__(256) void k_pair_resid(const float* __restrict__ Qp, const float* __restrict__ centroids,
                                                    const int* __restrict__ probes, const int* __restrict__ qn,
                                                    int64_t nitems, int nprobe, int r0, int nranks, int d, int dpad,
                                                    bf16x8* __restrict__ out, float* __restrict__ qerr) {
    // item = (query, rank in [r0, nranks)); its slot in the [nq][nprobe] arrays
    const int64_t it = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    const int nr = nranks - r0;
    const int64_t qi = hv_idiv(it, nr);
    const int j = r0 + (int)(it - qi * nr);
    const int64_t slot = qi * nprobe + j;
    const int g = threadIdx.x % G;
    const bool live = it < nitems && !(qn && j >= qn[qi]);
    const int k0 = g * 8;
    float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, se = 0.f, sr = 0.f;
    bf16x8 r;
    if (live) load8_diff(Qp + qi * d, centroids + (int64_t)probes[slot] * d, k0, d, v);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        r[e] = (__bf16)v[e];
        const float er = v[e] - (float)r[e];
        se = __builtin_fmaf(er, er, se);
        sr = __builtin_fmaf(v[e], v[e], sr);
    }
    for (int o = G / 2; o > 0; o >>= 1) {  // within the group
        se += __shfl_xor(se, o, 64);
        sr += __shfl_xor(sr, o, 64);
    }
    if (!live) return;
    if (k0 < dpad) out[slot * (dpad / 8) + g] = r;
    if (g == 0) qerr[slot] = __builtin_sqrtf(se) + 0x1p-23f * __builtin_sqrtf(sr);
}

void launch_pair_resid(const float* Qp, const float* centroids, const int* probes, const int* qn, int64_t nq,
                       int nprobe, int r0, int nranks, int d, int dpad, __bf16* out, float* qerr, hipStream_t s) {
    const int64_t nitems = nq * (nranks - r0);
    if (nitems <= 0) return;
    const int words = dpad / 8;  // <= 64
#define HV_RESID(G)                                                                                          \
    hipLaunchKernelGGL(k_pair_resid<G>, dim3((unsigned)((nitems + (256 / G) - 1) / (256 / G))), dim3(256), 0, s, \
                       Qp, centroids, probes, qn, nitems, nprobe, r0, nranks, d, dpad, reinterpret_cast<bf16x8*>(out), qerr)
    if (words <= 4)
        HV_RESID(4);
    else if (words <= 8)
        HV_RESID(8);
    else if (words <= 16)
        HV_RESID(16);
    else if (words <= 32)
        HV_RESID(32);
    else
        HV_RESID(64);
#undef HV_RESID
}

// FP8 codes of every list from the rows in list order: one workgroup per list finds the largest
// residual magnitude, then writes the scaled e4m3 codes in the BF16 codes' order (one byte per
// value) and the list's inverse scale.
__global__ __launch_bounds__(256) void k_build_resid_fp8(const float* __restrict__ rows, int d, int dpad,
                                                         const int64_t* __restrict__ list_off,
                                                         const float* __restrict__ centroids,
                                                         const int64_t* __restrict__ roff,
                                                         uint8_t* __restrict__ codes, float* __restrict__ inv_scale) {
    __shared__ float s_max[4];
    const int l = blockIdx.x;
    const int64_t b = list_off[l];
    const int n = (int)(list_off[l + 1] - b);
    const int n16 = (n + 15) / 16 * 16;
    const float* __restrict__ c = centroids + (int64_t)l * d;
    float m = 0.f;
    for (int64_t e = threadIdx.x; e < (int64_t)n * d; e += blockDim.x) {
        const int v = (int)(e / d), k = (int)(e - (int64_t)v * d);
        m = fmaxf(m, fabsf(rows[(b + v) * d + k] - c[k]));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    const float inv = fp8_inv_scale(fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3])));
    const float sc = 1.f / inv;  // a power of two
    if (threadIdx.x == 0) inv_scale[l] = inv;
    uint8_t* __restrict__ out = codes + roff[l];
    const int nks = dpad / 32;
    // one thread per 8 consecutive dims of a vector (one packed code word)
    const int g8 = dpad / 8;
    for (int64_t e = threadIdx.x; e < (int64_t)n16 * g8; e += blockDim.x) {
        const int v = (int)(e / g8), k0 = (int)(e - (int64_t)v * g8) * 8;
        float r[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = k0 + t;
            r[t] = (v < n && k < d) ? (rows[(b + v) * d + k] - c[k]) * sc : 0.f;
        }
        const int64_t o = (((int64_t)(v >> 4) * nks + (k0 >> 5)) * 64 + ((k0 & 31) >> 3) * 16 + (v & 15)) << 3;
        *reinterpret_cast<long*>(out + o) = pack_fp8x8(r);
    }
}

void launch_build_resid_fp8(const float* rows_by_pos, int d, int dpad, const int64_t* list_off,
                            const float* centroids, const int64_t* roff, int nlist, uint8_t* codes, float* inv_scale,
                            hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_build_resid_fp8, dim3((unsigned)nlist), dim3(256), 0, s, rows_by_pos, d, dpad, list_off,
                       centroids, roff, codes, inv_scale);
}

// FP8 query residuals of every (query, probe rank) slot: a group of G lanes per slot (G = the
// slot's 8-dim words rounded up to a power of two, <= 64; lane g holds dims 8g .. 8g+7), scaled
// by the slot's power of two.
template <int G>
__global__ __launch_bounds__(256) void k_pair_resid_fp8(const float* __restrict__ Qp,
                                                        const float* __restrict__ centroids,
                                                        const int* __restrict__ probes, const int* __restrict__ qn,
                                                        int64_t nitems, int nprobe, int nranks, int d, int dpad,
                                                        long* __restrict__ out, float* __restrict__ inv_scale) {
    const int64_t it = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;  // (query, rank < nranks)
    const int64_t qi = hv_idiv(it, nranks);
    const int j = (int)(it - qi * nranks);
    const int64_t slot = qi * nprobe + j;
    const int g = threadIdx.x % G;
    const bool live = it < nitems && !(qn && j >= qn[qi]);
    float r[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, m = 0.f;
    const int k0 = g * 8;
    if (live) {
        load8_diff(Qp + qi * d, centroids + (int64_t)probes[slot] * d, k0, d, r);
#pragma unroll
        for (int t = 0; t < 8; ++t) m = fmaxf(m, fabsf(r[t]));
    }
    for (int o = G / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));  // within the group
    if (!live) return;
    const float inv = fp8_inv_scale(m);
    const float sc = 1.f / inv;
#pragma unroll
    for (int t = 0; t < 8; ++t) r[t] *= sc;
    if (k0 < dpad) out[slot * (dpad / 8) + g] = pack_fp8x8(r);
    if (g == 0) inv_scale[slot] = inv;
}

void launch_pair_resid_fp8(const float* Qp, const float* centroids, const int* probes, const int* qn, int64_t nq,
                           int nprobe, int nranks, int d, int dpad, uint8_t* out, float* inv_scale, hipStream_t s) {
    const int64_t nitems = nq * nranks;
    if (nitems <= 0) return;
    const int words = dpad / 8;  // <= 64
#define HV_RESID8(G)                                                                                           \
    hipLaunchKernelGGL(k_pair_resid_fp8<G>, dim3((unsigned)((nitems + (256 / G) - 1) / (256 / G))), dim3(256), 0, s, \
                       Qp, centroids, probes, qn, nitems, nprobe, nranks, d, dpad, reinterpret_cast<long*>(out),   \
                       inv_scale)
    if (words <= 4)
        HV_RESID8(4);
    else if (words <= 8)
        HV_RESID8(8);
    else if (words <= 16)
        HV_RESID8(16);
    else if (words <= 32)
        HV_RESID8(32);
    else
        HV_RESID8(64);
#undef HV_RESID8
}

// ---------------------------------------------------------------------------------------
// The BF16 filter scan. Same work items as the FP32 scan (list, <= 64 queries, vector range),
// assigned statically: workgroup b takes the items j, j + G/8, j + 2G/8, ... of the queue of its
// XCD (b % 8), j = b / 8, G = grid size (the queues deal lists round-robin by size, so each
// stride of items is similar work). The item's queries form up to 4 blocks of 16. A wavefront
// holds the residuals of QBM query blocks in registers (NKS K-steps of 8 values per lane and
// block) and streams 16-vector blocks (the next one in flight), multiplying each into every
// block it holds; the wavefronts of the workgroup split the query blocks, and the vector blocks
// when there are fewer query blocks than wavefronts. QBM = 1: holding more query blocks saves
// code loads but the registers it takes cost more occupancy than that saves (measured at d = 128:
// the remaining-lists phase 0.77 ms with 4, 0.64 with 2, 0.60 with 1). After the MFMAs each lane
// tests the 4 (query, vector) pairs of each block it holds. Keys carry the vector's position in
// the list-ordered arrays, not its id.
//   mode 0: append the pairs with lb <= bound to the query's buffer, key = lb.
//   mode 1: the query's nearest list (probe rank 0 only): every pair of vector v is written to
//           slot v (no atomics; the caller limits the vectors to v < cap and sets the counts with
//           k_first_counts), key = lb, aux = ub.
// qaux: per slot, the FP8 inverse scale or the BF16 rounding-error norm |eq|; rerr: per vector,
// |ex| (BF16 only).
template <int NKS, typename E>
__global__ __launch_bounds__(256) void k_mfma_scan(const typename E::elem* __restrict__ qres,
                                                   const float* __restrict__ qaux,
                                                   const float* __restrict__ pdist, int nprobe,
                                                   const typename E::elem* __restrict__ rcodes,
                                                   const float* __restrict__ lscale,
                                                   const int64_t* __restrict__ roff,
                                                   const float* __restrict__ rnorm,
                                                   const float* __restrict__ rerr,
                                                   const int64_t* __restrict__ list_off,
                                                   const float* __restrict__ bound,
                                                   const int* __restrict__ pair_slot,
                                                   const ScanItem* __restrict__ items, const int* __restrict__ queue,
                                                   int vmax, int vsplit, int mode, float kappa1, float kappa_g,
                                                   float kappa2,
                                                   unsigned long long* __restrict__ surv, float* __restrict__ aux,
                                                   unsigned int* __restrict__ surv_cnt, int cap,
                                                   unsigned long long* __restrict__ stats) {
    constexpr int QBM = 1;  // query blocks per wavefront
    constexpr bool kPrefetch = NKS <= 8;                     // larger K keeps occupancy instead
    __shared__ int s_q[64], s_slot[64];
    // per query row of the item: |rq|^2, the bound (NaN past the item's queries: every
    // comparison fails), and the row's factors of the estimate and its error bound
    __shared__ __align__(16) float s_nrq[64];
    __shared__ __align__(16) float s_srq[64];  // BF16: |rq| + |eq|
    __shared__ __align__(16) float s_bound[64];
    __shared__ __align__(16) float s_qsc[64];  // BF16: kappa1 |eq|; FP8: -2 x the list's and slot's scales
    typedef typename E::vec vec;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col16 = lane & 15;
    const int xcd = blockIdx.x % kNumXcd;
    const int per_xcd = max(1, (int)gridDim.x / kNumXcd);
    const int qlen = queue[xcd], qbase = queue[2 * kNumXcd + xcd];
    unsigned long long pairs_done = 0;

    for (int slot_i = (int)blockIdx.x / kNumXcd; slot_i < qlen; slot_i += per_xcd) {
        const ScanItem item = items[qbase + slot_i];
        const int nQ = item.pair_end - item.pair_begin;
        const int l = item.list;
        if (tid < 64) {
            int q = 0, slot = 0;
            float nr = 0.f, b = __builtin_nanf(""), qa = 0.f;
            if (tid < nQ) {
                slot = pair_slot[item.pair_begin + tid];
                q = slot / nprobe;
                nr = pdist[slot];
                b = mode == 1 ? __builtin_inff() : bound[q];  // mode 1 keeps every pair
                qa = qaux[slot];
            }
            s_q[tid] = q;
            s_slot[tid] = slot;
            s_nrq[tid] = nr;
            s_srq[tid] = __builtin_amdgcn_sqrtf(nr) + (E::kScaled ? 0.f : qa);
            s_bound[tid] = b;
            s_qsc[tid] = E::kScaled ? -2.f * lscale[l] * qa : kappa1 * qa;  // powers of two: exact
        }
        const int64_t lbeg = list_off[l];
        const int nl = (int)min(list_off[l + 1] - lbeg, (long long)vmax);
        const int vend = min(nl, item.vec_begin + vsplit);
        __syncthreads();

        const int QB = (nQ + 15) / 16;
        const int ngrp = (QB + QBM - 1) / QBM;  // query-block groups: 1, 2, 3 or 4
        const int ngp = ngrp == 1 ? 1 : (ngrp == 2 ? 2 : 4);
        const int wpg = 4 / ngp;  // wavefronts per group, splitting its vector blocks
        const int grp = w / wpg, vsub = w % wpg;
        if (grp < ngrp && nl > 0) {
            const int qb0 = grp * QBM;
            const int nqb = min(QBM, QB - qb0);
            // A operands: query row (qb0 + j)*16 + col16, dims s*32 + g*8 .. +8. Rows past nQ read
            // slot s_slot[0]; their bound is -1.
            vec a[QBM][NKS];
#pragma unroll
            for (int j = 0; j < QBM; ++j) {
                const int qrow = (qb0 + j) * 16 + col16;
                const vec* __restrict__ qa = reinterpret_cast<const vec*>(qres) +
                                                (int64_t)s_slot[qrow < nQ ? qrow : 0] * (NKS * 4) + g;
#pragma unroll
                for (int s = 0; s < NKS; ++s) a[j][s] = qa[s * 4];
            }
            const vec* __restrict__ R = reinterpret_cast<const vec*>(rcodes + roff[l]) + lane;
            const int vstep = 16 * wpg;
            int v0 = item.vec_begin + vsub * 16;
            // software pipeline: the codes and norm of the next block are in flight while the
            // current block is computed
            vec bn[NKS];
            float nxn = 0.f, exn = 0.f;
            if (kPrefetch && v0 < vend) {
                const int64_t blk = (int64_t)(v0 >> 4) * NKS * 64;
#pragma unroll
                for (int s = 0; s < NKS; ++s) bn[s] = R[blk + s * 64];
                nxn = rnorm[lbeg + min(v0 + col16, nl - 1)];
                if (!E::kScaled) exn = rerr[lbeg + min(v0 + col16, nl - 1)];
            }
            for (; v0 < vend; v0 += vstep) {
                vec bc[NKS];
                float nx, ex;
                const int v1 = kPrefetch ? v0 + vstep : v0;
                if (kPrefetch) {
#pragma unroll
                    for (int s = 0; s < NKS; ++s) bc[s] = bn[s];
                    nx = nxn;
                    ex = exn;
                }
                if (v1 < vend) {
                    const int64_t blk = (int64_t)(v1 >> 4) * NKS * 64;
#pragma unroll
                    for (int s = 0; s < NKS; ++s) bn[s] = R[blk + s * 64];
                    nxn = rnorm[lbeg + min(v1 + col16, nl - 1)];
                    if (!E::kScaled) exn = rerr[lbeg + min(v1 + col16, nl - 1)];
                }
                if (!kPrefetch) {
#pragma unroll
                    for (int s = 0; s < NKS; ++s) bc[s] = bn[s];
                    nx = nxn;
                    ex = exn;
                }
                floatx4 acc[QBM];
#pragma unroll
                for (int j = 0; j < QBM; ++j) {
                    acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
                    if (j < nqb) {
#pragma unroll
                        for (int s = 0; s < NKS; ++s)
                            acc[j] = E::mma(a[j][s], bc[s], acc[j]);
                    }
                }
                const int v = v0 + col16;
                const bool cvalid = v < vend;
                // per-vector factors of the error bound (BF16):
                //   err = kappa1 |eq| sx + (|rq| + |eq|) vx + kappa2 (|rq|^2 + |rx|^2),
                //   vx = kappa1 |ex| + kappa_g (sx + |ex|)
                const float sx = E::kScaled ? 0.f : __builtin_amdgcn_sqrtf(nx);  // within 1 ulp (covered)
                const float vx = E::kScaled ? 0.f : __builtin_fmaf(kappa1, ex, kappa_g * (sx + ex));
                const uint32_t pos = (uint32_t)(lbeg + v);
                pairs_done += 256 * nqb;
#pragma unroll
                for (int j = 0; j < QBM; ++j) {
                    if (j >= nqb) break;
                    int r0 = (qb0 + j) * 16 + g * 4;
                    // re-read the rows' state from LDS each block rather than holding all of it in
                    // registers across the loop (occupancy)
                    asm volatile("" : "+v"(r0));
                    const float4 nr4 = *reinterpret_cast<const float4*>(&s_nrq[r0]);
                    const float4 b4 = *reinterpret_cast<const float4*>(&s_bound[r0]);
                    const float4 q4 = *reinterpret_cast<const float4*>(&s_qsc[r0]);
                    const float nrq[4] = {nr4.x, nr4.y, nr4.z, nr4.w};
                    const float bq[4] = {b4.x, b4.y, b4.z, b4.w};
                    const float qa[4] = {q4.x, q4.y, q4.z, q4.w};
                    float sqp[4] = {0.f, 0.f, 0.f, 0.f};
                    if (!E::kScaled) {
                        const float4 sr4 = *reinterpret_cast<const float4*>(&s_srq[r0]);
                        sqp[0] = sr4.x;
                        sqp[1] = sr4.y;
                        sqp[2] = sr4.z;
                        sqp[3] = sr4.w;
                    }
                    // estimate dt and its error bound er (FP8: estimates only, no bound); the
                    // pair is kept when dt - er <= bound (never for rows past the item's queries)
                    float dtv[4], erv[4];
                    bool kp[4];
                    bool anyk = false;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float nsum = nrq[i] + nx;
                        if (E::kScaled) {
                            dtv[i] = __builtin_fmaf(acc[j][i], qa[i], nsum);
                            erv[i] = 0.f;
                        } else {
                            dtv[i] = __builtin_fmaf(-2.f, acc[j][i], nsum);
                            erv[i] = __builtin_fmaf(qa[i], sx, __builtin_fmaf(sqp[i], vx, kappa2 * nsum));
                        }
                        kp[i] = cvalid && dtv[i] - erv[i] <= bq[i];
                        anyk |= kp[i];
                    }
                    if (mode == 1) {
                        // direct slots: each query row has one list here
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            if (kp[i] && v < cap) {
                                const int64_t o = (int64_t)s_q[r0 + i] * cap + v;
                                surv[o] = hv_key(fmaxf(dtv[i] - erv[i], 0.f), pos);
                                aux[o] = fmaxf(dtv[i] + erv[i], 0.f);
                            }
                        }
                        continue;
                    }
                    if (__ballot(anyk) == 0ull) continue;
                    // appends: one atomic per (query row, 16-lane group), all four issued before use
                    unsigned int gm[4], base[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        gm[i] = (unsigned int)((__ballot(kp[i]) >> (g * 16)) & 0xffffu);
                        base[i] = 0u;
                    }
                    if (col16 == 0) {
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            if (gm[i])
                                base[i] = atomicAdd(&surv_cnt[s_q[r0 + i]], (unsigned int)__builtin_popcount(gm[i]));
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const unsigned int bs = (unsigned int)__shfl((int)base[i], g * 16, 64);
                        if (kp[i]) {
                            const unsigned int slot = bs + (unsigned int)__builtin_popcount(gm[i] & ((1u << col16) - 1u));
                            if (slot < (unsigned int)cap) {
                                const int64_t o = (int64_t)s_q[r0 + i] * cap + slot;
                                surv[o] = hv_key(fmaxf(dtv[i] - erv[i], 0.f), pos);
                                aux[o] = fmaxf(dtv[i] + erv[i], 0.f);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if ((tid & 63) == 0 && pairs_done) atomicAdd(&stats[3], pairs_done);
}

// Candidate counts after the direct-slot scan of the nearest lists (vectors v < cap).
__global__ void k_first_counts(const int* __restrict__ probes, int64_t nq, int nprobe,
                               const int64_t* __restrict__ list_off, int cap, unsigned int* __restrict__ cnt) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int l = probes[q * nprobe];
    cnt[q] = (unsigned int)min(list_off[l + 1] - list_off[l], (long long)cap);
}

void launch_first_counts(const int* probes, int64_t nq, int nprobe, const int64_t* list_off, int cap,
                         unsigned int* cnt, hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_first_counts, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, probes, nq, nprobe,
                       list_off, cap, cnt);
}

int bf16_ksteps(int d) { return (d + 31) / 32; }

void launch_mfma_scan(int grid, int elem, const void* qres, const float* qaux, int d, const float* pdist,
                      int nprobe, const void* rcodes, const float* lscale, const int64_t* roff, const float* rnorm,
                      const float* rerr, const int64_t* list_off, const float* bound, const int* pair_slot,
                      const ScanItem* items, const int* queue, int vmax, int vsplit, int mode, float kappa1,
                      float kappa_g, float kappa2, unsigned long long* surv, float* aux, unsigned int* surv_cnt,
                      int cap, unsigned long long* stats, hipStream_t s) {
#define HV_SCAN_LAUNCH(K, E)                                                                                     \
    hipLaunchKernelGGL((k_mfma_scan<K, E>), dim3((unsigned)grid), dim3(256), 0, s,                               \
                       reinterpret_cast<const E::elem*>(qres), qaux, pdist, nprobe,                             \
                       reinterpret_cast<const E::elem*>(rcodes), lscale, roff, rnorm, rerr, list_off, bound,    \
                       pair_slot, items, queue, vmax, vsplit, mode, kappa1, kappa_g, kappa2, surv, aux, surv_cnt, \
                       cap, stats)
#define HV_SCAN_CASE(K)                \
    case K:                            \
        if (elem == 1)                 \
            HV_SCAN_LAUNCH(K, ElemFp8);  \
        else                           \
            HV_SCAN_LAUNCH(K, ElemBf16); \
        break;
    switch (bf16_ksteps(d)) {
        HV_SCAN_CASE(1)
        HV_SCAN_CASE(2)
        HV_SCAN_CASE(3)
        HV_SCAN_CASE(4)
        HV_SCAN_CASE(5)
        HV_SCAN_CASE(6)
        HV_SCAN_CASE(7)
        HV_SCAN_CASE(8)
        HV_SCAN_CASE(9)
        HV_SCAN_CASE(10)
        HV_SCAN_CASE(11)
        HV_SCAN_CASE(12)
        HV_SCAN_CASE(13)
        HV_SCAN_CASE(14)
        HV_SCAN_CASE(15)
        HV_SCAN_CASE(16)
        default: break;
    }
#undef HV_SCAN_CASE
#undef HV_SCAN_LAUNCH
}

int bf16_grid_size(int device) {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mfma_scan<4, ElemBf16>, 256, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    // the static item assignment deals the items of each XCD queue over grid / 8 workgroups
    return std::max(kNumXcd, ncu * per_cu / kNumXcd * kNumXcd);
}

// ---------------------------------------------------------------------------------------
// The BF16 form's nearest-list pass: each query's nearest list scanned and reduced in one kernel,
// in place of the direct-slot scan of the nearest lists (every pair written out) and the
// compaction that reads them back. A workgroup takes an item (list l, up to 16 of the queries
// whose nearest list l is, a range of up to kNearSplit vectors); each wavefront sweeps blocks w,
// w + 4, ... of the range with the filter scan's arithmetic (the queries' residuals in registers,
// the codes of its next two blocks in flight) and writes every pair's lower and upper bound into a
// fixed LDS slot, as 16-bit fixed-point values on the row's scale (lower bounds rounded down, upper
// bounds up, so both stay bounds; the scale spans the largest distance the list allows, (|q - c| +
// the list's radius)^2, to a few units of resolution). The sweep has no atomics and no barriers.
// Then each query's wavefront finds the k-th smallest upper bound of the range exactly, and the
// query's bound drops to it (when lower); the pairs with lower bound <= the bound are written to
// the query's survivor buffer. A list held by one item writes them from slot 0 with the count, the
// bound and the compaction count (`last`): what the nearest-list scan and its compaction leave. A
// list split into several items appends (atomics on the count), lowers the bound by an atomic
// minimum and leaves `last`, so the compaction after the pass merges the parts (each part keeps
// every pair of its range whose upper bound is among the k smallest of the list). A fused pass
// (`fused`: the items of several probe ranks in one launch, so a query's items may run at once)
// appends every row the same way; the k-th smallest upper bound over any part of a query's
// candidates bounds its k-th distance, so each part may lower the bound. There the nearest list of
// a query held by one item records its count in `last`, and the compaction leaves the query alone
// when nothing else was appended.
constexpr int kNearRows = kNearQueries;  // queries per item: one MFMA query block
static_assert(kNearRows == 16, "an item's queries are one 16-row MFMA block");
constexpr int kNearStride = kNearSplit + 16;  // LDS words per row (padded: the rows' banks differ)
constexpr unsigned int kNearTop = 65535u;     // code of "at or above the scale's end"
constexpr int kNearHist = 256 + 64;           // LDS words per wavefront: 256 bins, a word per lane
constexpr unsigned int kNearStage = 256u;     // kept positions a row stages in those words

// The r-th smallest (1-based, r <= the number of values) of the 16-bit values a wavefront holds E
// per lane, exact. The values are the low halves of packed words (0xffffffff: no value, not
// counted). A radix select over the high byte, then the low byte of the values in the selected
// high-byte bin, each pass a 256-bin histogram in LDS (hist: kNearHist words of this wavefront's
// own, bins 0..255 and one word per lane past them) built with atomic adds and read back with one
// prefix scan over the lanes (4 bins per lane). Every lane adds for every value, a value that is
// not counted into its lane's own word past the bins, so the adds need no change of the
// execution mask. All 64 lanes must call it.
template <int E>
__device__ __forceinline__ unsigned int wave_kth_u16(const unsigned int (&w)[E], int r, unsigned int* hist) {
    const int lane = threadIdx.x & 63;
    unsigned int prefix = 0u;  // the selected high byte (pass 2 counts only values with it)
    int rr = r;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        *reinterpret_cast<uint4*>(hist + 4 * lane) = make_uint4(0u, 0u, 0u, 0u);
        lds_fence();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const unsigned int v = w[e] & 0xffffu;
            const bool in = w[e] != 0xffffffffu && (pass == 0 || (v >> 8) == prefix);
            atomicAdd(&hist[in ? (pass == 0 ? v >> 8 : v & 255u) : 256u + (unsigned int)lane], 1u);
        }
        lds_fence();
        const uint4 h4 = *reinterpret_cast<const uint4*>(hist + 4 * lane);
        const unsigned int h[4] = {h4.x, h4.y, h4.z, h4.w};
        const unsigned int sum = h[0] + h[1] + h[2] + h[3];
        const unsigned int incl = wave_incl_scan(sum);
        const unsigned int excl = incl - sum;
        const bool mine = excl < (unsigned int)rr && (unsigned int)rr <= incl;
        unsigned int bin = 0u, below = 0u;
        if (mine) {
            unsigned int run = excl;
            bin = 4u * lane + 3u;
            below = run + h[0] + h[1] + h[2];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (run + h[t] >= (unsigned int)rr) {
                    bin = 4u * lane + t;
                    below = run;
                    break;
                }
                run += h[t];
            }
        }
        const int src = __builtin_ctzll(__ballot(mine) | (1ull << 63));
        bin = (unsigned int)__builtin_amdgcn_readlane((int)bin, src);
        below = (unsigned int)__builtin_amdgcn_readlane((int)below, src);
        rr -= (int)below;
        prefix = pass == 0 ? bin : (prefix << 8 | bin);
        lds_fence();  // the histogram is cleared again only after every lane has read it
    }
    return prefix;
}

template <int NKS>
__global__ __launch_bounds__(256) void k_mfma_nearest(const float* __restrict__ Qp,
                                                      const float* __restrict__ centroids, int d,
                                                      const float* __restrict__ pdist, int nprobe,
                                                      const __bf16* __restrict__ rcodes,
                                                      const int64_t* __restrict__ roff,
                                                      const float* __restrict__ rnorm,
                                                      const float* __restrict__ rerr,
                                                      const float* __restrict__ lrad,
                                                      const int64_t* __restrict__ list_off,
                                                      float* __restrict__ bound, float* __restrict__ bound_ref,
                                                      const int* __restrict__ pair_slot,
                                                      const ScanItem* __restrict__ items,
                                                      const int* __restrict__ queue, float kappa1, float kappa_g,
                                                      float kappa2, int k_all, const int* __restrict__ kq,
                                                      unsigned long long* __restrict__ surv, float* __restrict__ aux,
                                                      unsigned int* __restrict__ surv_cnt,
                                                      unsigned int* __restrict__ last, int cap, int fused,
                                                      unsigned long long* __restrict__ stats) {
    typedef ElemBf16::vec vec;
    __shared__ int s_q[kNearRows], s_kr[kNearRows], s_first[kNearRows];
    __shared__ __align__(16) float s_nrq[kNearRows];   // |rq|^2
    __shared__ __align__(16) float s_srq[kNearRows];   // |rq| + |eq|
    __shared__ __align__(16) float s_qsc[kNearRows];   // kappa1 |eq|
    __shared__ __align__(16) float s_inv[kNearRows];   // 1 / the row's fixed-point step
    __shared__ float s_step[kNearRows], s_b0[kNearRows];
    __shared__ unsigned int s_code[kNearRows][kNearStride];  // lower bound << 16 | upper bound
    __shared__ __align__(16) unsigned int s_hist[4][kNearHist];  // each wavefront's radix histogram
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = lane >> 4, col16 = lane & 15;
    const int xcd = blockIdx.x % kNumXcd;
    const int per_xcd = max(1, (int)gridDim.x / kNumXcd);
    const int qlen = queue[xcd], qbase = queue[2 * kNumXcd + xcd];
    unsigned long long pairs_done = 0;

    for (int slot_i = (int)blockIdx.x / kNumXcd; slot_i < qlen; slot_i += per_xcd) {
        const ScanItem item = items[qbase + slot_i];
        const int nQ = min(item.pair_end - item.pair_begin, kNearRows);
        const int l = item.list;
        const int64_t lbeg = list_off[l];
        const int nl_all = (int)(list_off[l + 1] - lbeg);
        const bool split = nl_all > kNearSplit;  // the list is cut into several items
        const int v_lo = item.vec_begin;
        const int nl = min(nl_all, v_lo + kNearSplit);  // end of the item's range
        const int nv = nl - v_lo;
        const int blk_lo = v_lo >> 4, nblk = ((nl + 15) >> 4) - blk_lo;
        const uint32_t pos0 = (uint32_t)(lbeg + v_lo);
        const vec* __restrict__ R = reinterpret_cast<const vec*>(rcodes + roff[l]) + lane;
        // the rows' state (lane r of wavefront 0 for query row r) and the A operands (query row
        // col16, dims s*32 + g*8 .. +8; rows past nQ read the first query's slot): the residual
        // r = q - c formed here in FP32 and split into BF16 parts hi + lo (kSplit: r - hi is exact
        // in FP32 and lo is its BF16 rounding, so what is left, r - hi - lo, is about 2^-16 |r|,
        // and both parts are multiplied into the codes), else its BF16 rounding alone; qe = the
        // norm of what the parts leave out, plus 2^-23 |r| for the FP32 subtraction (the four
        // lanes of a row sum theirs)
        constexpr bool kSplit = NKS <= 8;
        int q = 0, kr = 0, first = 0;
        float nr = 0.f, b = 0.f, qe = 0.f;
        if (tid < nQ) {
            const int slot = pair_slot[item.pair_begin + tid];
            q = slot / nprobe;
            first = slot - q * nprobe == 0;  // the query's nearest list
            nr = pdist[slot];
            b = bound[q];
            kr = hv_query_k(kq, q, k_all);
        }
        const float rad = lrad[l];
        vec a[NKS], alo[kSplit ? NKS : 1];
        {
            const int slot = pair_slot[item.pair_begin + (col16 < nQ ? col16 : 0)];
            const float* __restrict__ qv = Qp + (int64_t)(slot / nprobe) * d;
            const float* __restrict__ cv = centroids + (int64_t)l * d;
            float se = 0.f, sr = 0.f;
#pragma unroll
            for (int s = 0; s < NKS; ++s) {
                float v[8];
                load8_diff(qv, cv, s * 32 + g * 8, d, v);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    a[s][e] = (__bf16)v[e];
                    float er = v[e] - (float)a[s][e];
                    if (kSplit) {
                        alo[s][e] = (__bf16)er;
                        er -= (float)alo[s][e];
                    }
                    se = __builtin_fmaf(er, er, se);
                    sr = __builtin_fmaf(v[e], v[e], sr);
                }
            }
            se += __shfl_xor(se, 16, 64);
            sr += __shfl_xor(sr, 16, 64);
            se += __shfl_xor(se, 32, 64);
            sr += __shfl_xor(sr, 32, 64);
            if (tid < nQ) qe = __builtin_sqrtf(se) + 0x1p-23f * __builtin_sqrtf(sr);  // lane tid: row tid
        }
        // Wavefront w takes the range's blocks j = w, w + 4, w + 8, ..., two at a time, from two
        // buffers A and B used in turn (blocks 8t + w into A, 8t + 4 + w into B), each refilled
        // with its next block right after it is computed, so up to two blocks per wavefront are in
        // flight and no register copy waits for a load. Every refill is issued (past the range's
        // end as a repeat of its last block; an empty list reads the first list's first block), so
        // the number of loads behind each one is the same on every path and the waits for a block
        // are exact. Large K (more registers per block) keeps one buffer, loaded right before its
        // block.
        constexpr bool kTwo = NKS <= 4;
        vec bufA[NKS], bufB[kTwo ? NKS : 1];
        float nxA = 0.f, exA = 0.f, nxB = 0.f, exB = 0.f;
        const bool any_blk = nblk > 0;
        const vec* __restrict__ R0 = any_blk ? R : reinterpret_cast<const vec*>(rcodes) + lane;
        auto load_block = [&](int bk, vec* dst, float& nxd, float& exd) {
            bk = any_blk ? min(bk, nblk - 1) : -blk_lo;
            const int64_t blk = (int64_t)(blk_lo + bk) * NKS * 64;
#pragma unroll
            for (int s = 0; s < NKS; ++s) dst[s] = R0[blk + s * 64];
            const int64_t v = any_blk ? lbeg + min((blk_lo + bk) * 16 + col16, nl_all - 1) : 0;
            nxd = rnorm[v];
            exd = rerr[v];
        };
        if (kTwo) {
            load_block(w, bufA, nxA, exA);  // in flight during the item's set-up
            load_block(w + 4, bufB, nxB, exB);
        }
        if (tid < kNearRows) {
            // the row's scale: every exact distance of the list is <= (|rq| + radius)^2; 1/8 more
            // covers the estimate's error, a few ulps the FP32 rounding of the scale itself
            const float top = (__builtin_amdgcn_sqrtf(nr) + rad) * (__builtin_amdgcn_sqrtf(nr) + rad) * 1.125f + 1e-30f;
            const float step = top / 65000.f;
            s_q[tid] = q;
            s_kr[tid] = kr;
            s_first[tid] = first;
            s_nrq[tid] = nr;
            s_srq[tid] = __builtin_amdgcn_sqrtf(nr) + qe;
            s_qsc[tid] = kappa1 * qe;
            s_step[tid] = step;
            s_inv[tid] = 1.f / step;
            s_b0[tid] = b;
        }
        lds_barrier();  // the rows' state is in (the first blocks' loads stay in flight)

        // one 16-vector block: estimates and bounds of its 16 x 16 pairs, coded into LDS
        auto do_block = [&](int b0, const vec* bc, float nx, float ex) {
            floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NKS; ++s) {
                acc = ElemBf16::mma(a[s], bc[s], acc);
                if (kSplit) acc = ElemBf16::mma(alo[s], bc[s], acc);
            }
            const int v = (blk_lo + b0) * 16 + col16;
            // err = kappa1 |eq| sx + (|rq| + |eq|) vx + kappa2 (|rq|^2 + |rx|^2),
            // vx = kappa1 |ex| + kappa_g (sx + |ex|)
            // (with the split, 2d exact products are accumulated: the FP32 accumulation term
            // doubles, and 5 % more covers |hi| + |lo| against |r|)
            const float sx = __builtin_amdgcn_sqrtf(nx);  // within 1 ulp (covered)
            const float vx = __builtin_fmaf(kappa1, ex, (kSplit ? 2.1f : 1.f) * kappa_g * (sx + ex));
            pairs_done += 256;
            int r0 = g * 4;
            asm volatile("" : "+v"(r0));
            const float4 nr4 = *reinterpret_cast<const float4*>(&s_nrq[r0]);
            const float4 q4 = *reinterpret_cast<const float4*>(&s_qsc[r0]);
            const float4 sr4 = *reinterpret_cast<const float4*>(&s_srq[r0]);
            const float4 iv4 = *reinterpret_cast<const float4*>(&s_inv[r0]);
            const float nrq[4] = {nr4.x, nr4.y, nr4.z, nr4.w};
            const float qa[4] = {q4.x, q4.y, q4.z, q4.w};
            const float sqp[4] = {sr4.x, sr4.y, sr4.z, sr4.w};
            const float inv[4] = {iv4.x, iv4.y, iv4.z, iv4.w};
            const int slot = v - v_lo;  // < kNearSplit: the range's blocks stay inside it
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float nsum = nrq[i] + nx;
                const float dt = __builtin_fmaf(-2.f, acc[i], nsum);
                const float er = __builtin_fmaf(qa[i], sx, __builtin_fmaf(sqp[i], vx, kappa2 * nsum));
                // codes: floor(lb / step) - 2 (>= 0) and floor(ub / step) + 2: the margins of 2
                // steps cover the rounding of the products and of the decoding
                // (an upper bound at or past the scale's end, or NaN, codes as kNearTop: none)
                const float fl = fminf(fmaxf((dt - er) * inv[i], 0.f), 65000.f);
                const float fu = (dt + er) * inv[i];
                const unsigned int cl = (unsigned int)max((int)fl - 2, 0);
                const unsigned int cu = fu < 65000.f ? (unsigned int)fmaxf(fu, 0.f) + 2u : kNearTop;
                s_code[r0 + i][slot] = cl << 16 | cu;
            }
        };
        // everything the item loaded so far (the A operands, the first blocks) is in before the
        // sweep: the sweep then starts each step with exactly the two refills of the previous one
        // in flight, and the waits for a block cover its own loads only
        __builtin_amdgcn_s_waitcnt(0x0f70);  // s_waitcnt vmcnt(0)
        const int nsteps = (nblk + 7) >> 3;
        for (int t = 0; t < nsteps; ++t) {
            const int bA = 8 * t + w, bB = bA + 4;
            if (kTwo) {
                if (bA < nblk) do_block(bA, bufA, nxA, exA);
                load_block(bA + 8, bufA, nxA, exA);
                if (bB < nblk) do_block(bB, bufB, nxB, exB);
                load_block(bB + 8, bufB, nxB, exB);
            } else {
                if (bA < nblk) {
                    load_block(bA, bufA, nxA, exA);
                    do_block(bA, bufA, nxA, exA);
                }
                if (bB < nblk) {
                    load_block(bB, bufA, nxA, exA);
                    do_block(bB, bufA, nxA, exA);
                }
            }
        }
        lds_barrier();  // every pair's codes are in

        // per query row (wavefront w takes rows w, w + 4, ...): the exact k-th smallest upper
        // bound of the range, then the pairs with lower bound <= the query's bound
        constexpr int E = kNearSplit / 64;
        for (int r = w; r < nQ; r += 4) {
            const int qq = s_q[r], k_r = s_kr[r];
            const float step = s_step[r];
            float T = s_b0[r];
            unsigned int c[E];  // the pairs' packed codes (0xffffffff past the range)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int j = e * 64 + lane;
                c[e] = j < nv ? s_code[r][j] : 0xffffffffu;
            }
            if (nv >= k_r) {
                const unsigned int kth = wave_kth_u16<E>(c, k_r, s_hist[w]);
                if (kth < kNearTop) T = fminf(T, (float)kth * step);  // >= the k-th upper bound
            }
            // the kept pairs (lower bound <= T): each lane's kept positions staged, lane by lane,
            // in this wavefront's LDS words (s_hist[w], free after the select; a position past
            // the stage goes to the lane's own word past it, so every lane writes for every value
            // and the execution mask never changes), then written out coalesced; a row with more
            // than kNearStage kept pairs writes them directly instead
            // (a bitwise and of two tests: no short-circuit branch; past the range c = 0xffffffff)
            auto keep_of = [&](int e) { return (c[e] != 0xffffffffu) & ((float)(c[e] >> 16) * step <= T); };
            unsigned int* __restrict__ stage = s_hist[w];
            unsigned int kb = 0u;  // the lane's kept values, bit e for value e
#pragma unroll
            for (int e = 0; e < E; ++e) kb |= keep_of(e) ? 1u << e : 0u;
            const unsigned int kcnt = (unsigned int)__builtin_popcount(kb);
            const unsigned int kincl = wave_incl_scan(kcnt);
            const unsigned int m = (unsigned int)__builtin_amdgcn_readlane((int)kincl, 63);
            {
                unsigned int at_l = kincl - kcnt;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const unsigned int bit = (kb >> e) & 1u;
                    stage[bit && at_l < kNearStage ? at_l : kNearStage + (unsigned int)lane] = (unsigned int)(e * 64 + lane);
                    at_l += bit;
                }
            }
            unsigned int at = 0u;
            const bool append = split || fused;
            if (append) {
                // parts of one list, and every row of a fused pass, append (the compaction after
                // the pass merges them); in a fused pass the nearest list of a query held by one
                // item records its count, so a query nothing else appended to is left alone
                if (lane == 0) {
                    if (m) at = atomicAdd(&surv_cnt[qq], m);
                    atomicMin(reinterpret_cast<unsigned int*>(bound) + qq, __float_as_uint(T));
                    if (!split && s_first[r] && last) last[qq] = m;
                }
                at = (unsigned int)__builtin_amdgcn_readfirstlane((int)at);
            }
            unsigned long long* __restrict__ sq = surv + (int64_t)qq * cap;
            unsigned int* __restrict__ aq = reinterpret_cast<unsigned int*>(aux) + (int64_t)qq * cap;
            // one kept pair: its key (lower bound, position) and upper bound
            auto put = [&](unsigned int slot, unsigned int cc, unsigned int j) {
                if (slot < (unsigned int)cap) {
                    const unsigned int cu = cc & 0xffffu;
                    sq[slot] = hv_key((float)(cc >> 16) * step, pos0 + j);
                    aq[slot] = cu >= kNearTop ? 0x7f800000u : __float_as_uint((float)cu * step);
                }
            };
            if (m <= kNearStage) {
                lds_fence();  // the staged positions of every lane are in
                for (unsigned int i = (unsigned int)lane; i < m; i += 64u) {
                    const unsigned int j = stage[i];
                    put(at + i, s_code[r][j], j);
                }
            } else {
                unsigned int o = at;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const bool keep = keep_of(e);
                    const unsigned long long km = __ballot(keep);
                    if (keep)
                        put(o + __builtin_amdgcn_mbcnt_hi((unsigned int)(km >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned int)km, 0u)),
                            c[e], (unsigned int)(e * 64 + lane));
                    o += (unsigned int)__builtin_popcountll(km);
                }
            }
            lds_fence();  // the staging words are read before the next row's select clears them
            if (!append && lane == 0) {
                surv_cnt[qq] = m;
                if (last) last[qq] = m;
                bound[qq] = T;
                if (bound_ref) bound_ref[qq] = T;
            }
        }
        lds_barrier();  // the rows are free for the next item (the stores need not wait)
    }
    if (lane == 0 && pairs_done) atomicAdd(&stats[3], pairs_done);
}

void launch_mfma_nearest(int grid, const float* Qp, const float* centroids, int d, const float* pdist, int nprobe,
                         const __bf16* rcodes, const int64_t* roff, const float* rnorm, const float* rerr,
                         const float* lrad, const int64_t* list_off,
                         float* bound, float* bound_ref, const int* pair_slot, const ScanItem* items,
                         const int* queue, float kappa1, float kappa_g, float kappa2, int k, const int* kq,
                         unsigned long long* surv, float* aux, unsigned int* surv_cnt, unsigned int* last, int cap,
                         int fused, unsigned long long* stats, hipStream_t s) {
#define HV_NEAR_CASE(K)                                                                                              \
    case K:                                                                                                          \
        hipLaunchKernelGGL((k_mfma_nearest<K>), dim3((unsigned)grid), dim3(256), 0, s, Qp, centroids, d,             \
                           pdist, nprobe, rcodes, roff, rnorm, rerr, lrad, list_off, bound, bound_ref, pair_slot,    \
                           items, queue, kappa1, kappa_g, kappa2, k, kq, surv, aux, surv_cnt, last, cap, fused,      \
                           stats);                                                                                   \
        break;
    switch (bf16_ksteps(d)) {
        HV_NEAR_CASE(1)
        HV_NEAR_CASE(2)
        HV_NEAR_CASE(3)
        HV_NEAR_CASE(4)
        HV_NEAR_CASE(5)
        HV_NEAR_CASE(6)
        HV_NEAR_CASE(7)
        HV_NEAR_CASE(8)
        HV_NEAR_CASE(9)
        HV_NEAR_CASE(10)
        HV_NEAR_CASE(11)
        HV_NEAR_CASE(12)
        HV_NEAR_CASE(13)
        HV_NEAR_CASE(14)
        HV_NEAR_CASE(15)
        HV_NEAR_CASE(16)
        default: break;
    }
#undef HV_NEAR_CASE
}

// ---------------------------------------------------------------------------------------
// After the BF16 scan of each query's nearest list: the k-th smallest upper bound bounds the
// k-th exact distance, so the query's bound drops to it when tighter; the stored pairs whose
// lower bound exceeds the new bound cannot be results and are dropped (in-place compaction, order
// not kept). One workgroup per query holds the query's upper bounds and keys in registers
// (kCompactPer per thread, so buffers up to 256 * kCompactPer pairs), finds the k-th upper bound
// by a radix select (8-bit digits of the FP32 bits, which order non-negative floats) over a
// 256-bin LDS histogram, starting at the first byte where the smallest and largest values differ
// (all values share the bytes above it). An overflowed query (count > cap) keeps its count: its
// bound still tightens (the k-th over the stored subset is >= the k-th over all), and the exact
// path redoes it. With `last` given, a query whose count is what the previous compaction left
// (no pair appended since) is skipped: its buffer and bound would come out unchanged.
constexpr int kCompactPer = kMaxSelectRow / 256;
constexpr int kCompactWaveMax = 2048;  // one wavefront per query up to here with <= 32 values per lane,
constexpr int kCompactWaveBig = 4096;  // then with 64 per lane (same kernel); larger: the workgroup kernel

__global__ __launch_bounds__(256) void k_bf16_compact(unsigned long long* __restrict__ surv,
                                                      float* __restrict__ aux,
                                                      unsigned int* __restrict__ cnt, int64_t nq, int cap, int k_all,
                                                      float* __restrict__ bound, float* __restrict__ bound_ref,
                                                      unsigned int* __restrict__ last,
                                                      const int* __restrict__ kq, int kmul) {
    __shared__ unsigned int s_hist[256];
    __shared__ unsigned int s_red[2][4];
    __shared__ unsigned int s_sel[2];
    __shared__ unsigned int s_out;
    const int64_t q = blockIdx.x;
    const int k = kq ? min(k_all, kmul * hv_query_k(kq, q, k_all)) : k_all;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const unsigned int c = cnt[q];
    const int n = (int)min(c, (unsigned int)cap);
    if (n <= kCompactWaveBig || (last && last[q] == c)) return;  // k_bf16_compact_wave; unchanged
    float b = bound[q];
    unsigned long long* __restrict__ buf = surv + q * cap;
    const unsigned int* __restrict__ ubits = reinterpret_cast<const unsigned int*>(aux + q * cap);
    unsigned int* ubw = reinterpret_cast<unsigned int*>(aux + q * cap);
    unsigned int u[kCompactPer];
    unsigned long long key[kCompactPer];
#pragma unroll
    for (int i = 0; i < kCompactPer; ++i) {
        const int j = t + 256 * i;
        unsigned int v = 0xffffffffu;
        key[i] = HV_KEY_PAD;
        if (j < n) {
            v = ubits[j];
            v = (int)v < 0 ? 0u : v;  // -0.0 orders as 0
            key[i] = buf[j];
        }
        u[i] = v;
    }
    if (n >= k) {
        unsigned int lo = 0xffffffffu, hi = 0u;
#pragma unroll
        for (int i = 0; i < kCompactPer; ++i)
            if (t + 256 * i < n) {
                lo = min(lo, u[i]);
                hi = max(hi, u[i]);
            }
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, (unsigned int)__shfl_xor((int)lo, o, 64));
            hi = max(hi, (unsigned int)__shfl_xor((int)hi, o, 64));
        }
        if (lane == 0) {
            s_red[0][w] = lo;
            s_red[1][w] = hi;
        }
        __syncthreads();
        lo = min(min(s_red[0][0], s_red[0][1]), min(s_red[0][2], s_red[0][3]));
        hi = max(max(s_red[1][0], s_red[1][1]), max(s_red[1][2], s_red[1][3]));
        const unsigned int diff = lo ^ hi;
        const int top = diff ? (31 - __builtin_clz(diff)) / 8 * 8 : -8;  // shift of the first differing byte
        unsigned int mask = top >= 24 ? 0u : (0xffffffffu << (top + 8));
        unsigned int prefix = lo & mask;
        int kk = k;  // rank (1-based) still to find inside the current prefix
        for (int shift = top; shift >= 0; shift -= 8) {
            s_hist[t] = 0u;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kCompactPer; ++i)
                if (t + 256 * i < n && (u[i] & mask) == prefix) atomicAdd(&s_hist[(u[i] >> shift) & 255u], 1u);
            __syncthreads();
            if (w == 0) {
                unsigned int h[4], sum = 0u;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    h[e] = s_hist[lane * 4 + e];
                    sum += h[e];
                }
                unsigned int incl = sum;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const unsigned int y = (unsigned int)__shfl_up((int)incl, o, 64);
                    if (lane >= o) incl += y;
                }
                const unsigned int excl = incl - sum;
                if (excl < (unsigned int)kk && (unsigned int)kk <= incl) {
                    unsigned int run = excl;
                    int digit = lane * 4 + 3;
                    unsigned int below = run + h[0] + h[1] + h[2];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (run + h[e] >= (unsigned int)kk) {
                            digit = lane * 4 + e;
                            below = run;
                            break;
                        }
                        run += h[e];
                    }
                    s_sel[0] = (unsigned int)digit;
                    s_sel[1] = below;
                }
            }
            __syncthreads();
            kk -= (int)s_sel[1];
            prefix |= s_sel[0] << shift;
            mask |= 255u << shift;
        }
        b = fminf(b, __uint_as_float(prefix));
    }
    if (c > (unsigned int)cap) {
        if (t == 0) {
            bound[q] = b;
            if (bound_ref) bound_ref[q] = b;
            if (last) last[q] = c;
        }
        return;
    }
    // every key is in registers: rewrite the buffer with the pairs whose lower bound is <= b
    if (t == 0) s_out = 0u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kCompactPer; ++i) {
        const bool keep = t + 256 * i < n && hv_key_dist(key[i]) <= b;
        const unsigned long long m = __ballot(keep);
        unsigned int base = 0u;
        if (lane == 0 && m) base = atomicAdd(&s_out, (unsigned int)__builtin_popcountll(m));
        base = (unsigned int)__shfl((int)base, 0, 64);
        if (keep) {
            const unsigned int o = base + __builtin_popcountll(m & ((1ull << lane) - 1ull));
            buf[o] = key[i];
            ubw[o] = u[i];  // upper bounds move with their keys
        }
    }
    __syncthreads();
    if (t == 0) {
        cnt[q] = s_out;
        bound[q] = b;
        if (bound_ref) bound_ref[q] = b;
        if (last) last[q] = s_out;
    }
}

// The same with one wavefront per query for buffers of at most kCompactWaveMax pairs (upper
// bounds in registers, the k-th from wave_kth_u32, keys streamed through registers after it).

template <int E>
__device__ __forceinline__ void bf16_compact_wave_e(unsigned long long* __restrict__ buf,
                                                    unsigned int* __restrict__ ubits, int n, unsigned int c,
                                                    int cap, int k, float b, unsigned int* __restrict__ cnt_q,
                                                    float* __restrict__ bound_q, float* __restrict__ bref_q,
                                                    unsigned int* __restrict__ last_q, unsigned int* s_tmp) {
    const int lane = threadIdx.x & 63;
    if (n >= k) {
        unsigned int u[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = e * 64 + lane;
            unsigned int x = 0xffffffffu;
            if (j < n) {
                x = ubits[j];
                x = (int)x < 0 ? 0u : x;  // -0.0 orders as 0
            }
            u[e] = x;
        }
        b = fminf(b, __uint_as_float(wave_kth_u32<E>(u, k, s_tmp)));
    }
    if (c > (unsigned int)cap) {
        if (lane == 0) {
            *bound_q = b;
            if (bref_q) *bref_q = b;
            if (last_q) *last_q = c;
        }
        return;
    }
    // the keys pass through registers H per lane at a time: every key of a part is loaded before
    // any of its kept keys is stored, and stores land at or below the key's own position
    constexpr int H = E > 32 ? 32 : E;
    unsigned int out = 0u;
#pragma unroll
    for (int h0 = 0; h0 < E; h0 += H) {
        unsigned long long key[H];
        unsigned int ub[H];
#pragma unroll
        for (int e = 0; e < H; ++e) {
            const int j = (h0 + e) * 64 + lane;
            key[e] = j < n ? buf[j] : HV_KEY_PAD;
            ub[e] = j < n ? ubits[j] : 0u;
        }
#pragma unroll
        for (int e = 0; e < H; ++e) {
            const bool keep = (h0 + e) * 64 + lane < n && hv_key_dist(key[e]) <= b;
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const unsigned int o = out + __builtin_popcountll(m & ((1ull << lane) - 1ull));
                buf[o] = key[e];
                ubits[o] = ub[e];  // upper bounds move with their keys (later compactions read them)
            }
            out += (unsigned int)__builtin_popcountll(m);
        }
    }
    if (lane == 0) {
        *cnt_q = out;
        *bound_q = b;
        if (bref_q) *bref_q = b;
        if (last_q) *last_q = out;
    }
}

__global__ __launch_bounds__(256) void k_bf16_compact_wave(unsigned long long* __restrict__ surv,
                                                           float* __restrict__ aux,
                                                           unsigned int* __restrict__ cnt, int64_t nq, int cap, int k_all,
                                                           float* __restrict__ bound, float* __restrict__ bound_ref,
                                                           unsigned int* __restrict__ last,
                                                           const int* __restrict__ kq, int kmul) {
    __shared__ unsigned int s_tmp[4][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= nq) return;
    const int k = kq ? min(k_all, kmul * hv_query_k(kq, q, k_all)) : k_all;
    const unsigned int c = cnt[q];
    const int n = (int)min(c, (unsigned int)cap);
    if (n > kCompactWaveBig) return;  // k_bf16_compact
    if (last && last[q] == c) return;  // unchanged since the previous compaction
    unsigned long long* buf = surv + q * cap;
    unsigned int* ubits = reinterpret_cast<unsigned int*>(aux + q * cap);
    const float b = bound[q];
    unsigned int* lq = last ? last + q : nullptr;
    if (n <= 256)
        bf16_compact_wave_e<4>(buf, ubits, n, c, cap, k, b, cnt + q, bound + q, bound_ref ? bound_ref + q : nullptr, lq, s_tmp[w]);
    else if (n <= 512)
        bf16_compact_wave_e<8>(buf, ubits, n, c, cap, k, b, cnt + q, bound + q, bound_ref ? bound_ref + q : nullptr, lq, s_tmp[w]);
    else if (n <= 1024)
        bf16_compact_wave_e<16>(buf, ubits, n, c, cap, k, b, cnt + q, bound + q, bound_ref ? bound_ref + q : nullptr, lq, s_tmp[w]);
    else if (n <= kCompactWaveMax)
        bf16_compact_wave_e<32>(buf, ubits, n, c, cap, k, b, cnt + q, bound + q, bound_ref ? bound_ref + q : nullptr, lq, s_tmp[w]);
    else
        bf16_compact_wave_e<kCompactWaveBig / 64>(buf, ubits, n, c, cap, k, b, cnt + q, bound + q, bound_ref ? bound_ref + q : nullptr, lq, s_tmp[w]);
}

void launch_bf16_compact(unsigned long long* surv, float* aux, unsigned int* cnt, int64_t nq, int cap, int k,
                         float* bound, unsigned int* last, const int* kq, int kmul, hipStream_t s, float* bound_ref) {
    if (nq <= 0) return;
    const dim3 wgrid((unsigned)((nq + 3) / 4));
    hipLaunchKernelGGL(k_bf16_compact_wave, wgrid, dim3(256), 0, s, surv, aux, cnt, nq, cap, k, bound, bound_ref, last,
                       kq, kmul);
    if (cap > kCompactWaveBig)
        hipLaunchKernelGGL(k_bf16_compact, dim3((unsigned)nq), dim3(256), 0, s, surv, aux, cnt, nq, cap, k, bound,
                           bound_ref, last, kq, kmul);
}

// FP8 shortlist: each query keeps its kr smallest estimates (keys' distance bits; every key tied
// with the kr-th stays too), compacted in place, and its filter bound drops to the kr-th
// estimate. One wavefront per query, buffers up to kCompactWaveBig pairs (the FP8 form's
// survivor buffers are no larger); overflowed queries are left to the exact path.
template <int E>
__device__ __forceinline__ void shortlist_wave_e(unsigned long long* __restrict__ buf, int n, int kr,
                                                 unsigned int* __restrict__ cnt_q, float* __restrict__ bound_q,
                                                 unsigned int* s_tmp) {
    const int lane = threadIdx.x & 63;
    unsigned int kth;
    {
        unsigned int u[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = e * 64 + lane;
            u[e] = j < n ? (unsigned int)(buf[j] >> 32) : 0xffffffffu;
        }
        kth = wave_kth_u32<E>(u, kr, s_tmp);
    }
    constexpr int H = E > 32 ? 32 : E;
    unsigned int out = 0u;
#pragma unroll
    for (int h0 = 0; h0 < E; h0 += H) {
        unsigned long long key[H];
#pragma unroll
        for (int e = 0; e < H; ++e) {
            const int j = (h0 + e) * 64 + lane;
            key[e] = j < n ? buf[j] : HV_KEY_PAD;
        }
#pragma unroll
        for (int e = 0; e < H; ++e) {
            const bool keep = (h0 + e) * 64 + lane < n && (unsigned int)(key[e] >> 32) <= kth;
            const unsigned long long m = __ballot(keep);
            if (keep) buf[out + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = key[e];
            out += (unsigned int)__builtin_popcountll(m);
        }
    }
    if (lane == 0) {
        *cnt_q = out;
        *bound_q = fminf(*bound_q, __uint_as_float(kth));
    }
}

template <bool BIG>
__global__ __launch_bounds__(256) void k_shortlist(unsigned long long* __restrict__ surv,
                                                   unsigned int* __restrict__ cnt, int64_t nq, int cap, int kr_all,
                                                   float* __restrict__ bound, const int* __restrict__ kq, int kmul) {
    __shared__ unsigned int s_tmp[4][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= nq) return;
    const int kr = kq ? min(kr_all, kmul * hv_query_k(kq, q, kr_all)) : kr_all;
    const unsigned int c = cnt[q];
    if (c > (unsigned int)cap || c <= (unsigned int)kr) return;
    const int n = (int)c;
    unsigned long long* buf = surv + q * cap;
    if (BIG) {
        if (n > kCompactWaveMax && n <= kCompactWaveBig)
            shortlist_wave_e<kCompactWaveBig / 64>(buf, n, kr, cnt + q, bound + q, s_tmp[w]);
    } else if (n <= 256) {
        shortlist_wave_e<4>(buf, n, kr, cnt + q, bound + q, s_tmp[w]);
    } else if (n <= 512) {
        shortlist_wave_e<8>(buf, n, kr, cnt + q, bound + q, s_tmp[w]);
    } else if (n <= 1024) {
        shortlist_wave_e<16>(buf, n, kr, cnt + q, bound + q, s_tmp[w]);
    } else if (n <= kCompactWaveMax) {
        shortlist_wave_e<32>(buf, n, kr, cnt + q, bound + q, s_tmp[w]);
    }
}

void launch_shortlist(unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int kr, float* bound,
                      const int* kq, int kmul, hipStream_t s) {
    if (nq <= 0) return;
    const dim3 grid((unsigned)((nq + 3) / 4));
    hipLaunchKernelGGL(k_shortlist<false>, grid, dim3(256), 0, s, surv, cnt, nq, cap, kr, bound, kq, kmul);
    if (cap > kCompactWaveMax)
        hipLaunchKernelGGL(k_shortlist<true>, grid, dim3(256), 0, s, surv, cnt, nq, cap, kr, bound, kq, kmul);
}

// ---------------------------------------------------------------------------------------
// Flat exact re-rank. A query's stored pairs are cut into work items of 64 (one pair per lane),
// so a query with many survivors spreads over many wavefronts instead of holding one. Each
// pair's exact FP32 distance (the plain path's blocked FMA chains in dimension order) is kept
// only when it does not exceed the query's bound: the bound is the k-th smallest upper bound of
// the nearest list (or lower), so it is at least the k-th exact distance and no top-k member
// is dropped. Each pair's slot of the kept buffer gets its key (distance, id), or padding when
// dropped; the final select compacts and sorts them.
//
// k_rerank_items: one thread per query; lists overflowed queries, counts the re-ranked pairs
// (stats[2]) and reserves its items with one atomic per workgroup. With the (list, rank)-sorted
// pairs given, the queries are taken in the order of their rank-0 pairs, i.e. grouped by nearest
// list: queries of one list share many survivor rows, and the re-rank below runs neighbouring
// items on one XCD, so those rows are read from HBM once into its L2.
__global__ __launch_bounds__(256) void k_rerank_items(const unsigned int* __restrict__ cnt, int64_t nq, int cap,
                                                      const int* __restrict__ pairs, const int* __restrict__ pair_slot,
                                                      const int* __restrict__ npairs, int nprobe,
                                                      int2* __restrict__ items, unsigned int* __restrict__ n_items,
                                                      int* __restrict__ overflow_list, int* __restrict__ overflow_cnt,
                                                      unsigned long long* __restrict__ stats) {
    __shared__ unsigned int s_wsum[4];
    __shared__ unsigned int s_pairs[4];
    __shared__ unsigned int s_base;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t e = (int64_t)blockIdx.x * 256 + t;
    int64_t q = -1;
    if (pairs) {
        if (e < (int64_t)*npairs && pair_slot[e] % nprobe == 0) q = pairs[e];
    } else if (e < nq) {
        q = e;
    }
    unsigned int c = 0u;
    if (q >= 0) {
        c = cnt[q];
        if (c > (unsigned int)cap) {
            if (overflow_list) overflow_list[atomicAdd(overflow_cnt, 1)] = (int)q;
            c = 0u;
        }
    }
    const unsigned int nch = (c + 63u) / 64u;
    unsigned int incl = nch, pc = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int y = (unsigned int)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += y;
        pc += (unsigned int)__shfl_xor((int)pc, o, 64);
    }
    if (lane == 63) s_wsum[w] = incl;
    if (lane == 0) s_pairs[w] = pc;
    __syncthreads();
    if (t == 0) {
        const unsigned int tot = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
        s_base = tot ? atomicAdd(n_items, tot) : 0u;
        const unsigned long long np = (unsigned long long)s_pairs[0] + s_pairs[1] + s_pairs[2] + s_pairs[3];
        if (np) atomicAdd(&stats[2], np);
    }
    __syncthreads();
    unsigned int base = s_base + incl - nch;
    for (int i = 0; i < w; ++i) base += s_wsum[i];
    for (unsigned int i = 0; i < nch; ++i)
        items[base + i] = make_int2((int)q, (int)(i * 64u | min(64u, c - i * 64u) << 16));
}

// One staging chunk (<= kChunkDims dims from lo) of a row: V4 loads 16 bytes at a time (a
// float4 starting at or past d only covers dims past the chunk, as d % 4 == 0).
template <bool V4>
__device__ __forceinline__ void load_chunk(const float* __restrict__ v, int lo, int d, float (&r)[kChunkDims]) {
    if (V4) {
#pragma unroll
        for (int t = 0; t < kChunkDims / 4; ++t) {
            const float4 x = *reinterpret_cast<const float4*>(v + min(lo + 4 * t, d - 4));
            r[4 * t] = x.x;
            r[4 * t + 1] = x.y;
            r[4 * t + 2] = x.z;
            r[4 * t + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < kChunkDims; ++t) r[t] = v[min(lo + t, d - 1)];
    }
}

// p += (q - x)^2 over the chunk's len dims, in dimension order
template <bool V4>
__device__ __forceinline__ void chain_chunk(const float* __restrict__ qv, int lo, int len, int d,
                                            const float (&x)[kChunkDims], float& p) {
    float qr[kChunkDims];
    load_chunk<V4>(qv, lo, d, qr);
    if (len >= kChunkDims) {
        // a full chunk (len is wavefront-uniform): no per-dimension test
#pragma unroll
        for (int t = 0; t < kChunkDims; ++t) {
            const float u = qr[t] - x[t];
            p = __builtin_fmaf(u, u, p);
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < kChunkDims; ++t) {
        if (t < len) {
            const float u = qr[t] - x[t];
            p = __builtin_fmaf(u, u, p);
        }
    }
}

// The items are dealt to the XCDs in contiguous eighths (workgroup b runs on XCD b % 8; the
// grid is a multiple of 8), each XCD's wavefronts striding through its eighth, so items next to
// each other (the queries of one nearest list) run together under one L2.
template <bool V4>
__global__ __launch_bounds__(256) void k_rerank_flat(const float* __restrict__ Qp, int d,
                                                     const float* __restrict__ rows_by_pos,
                                                     const uint32_t* __restrict__ ids, ChunkPlan plan,
                                                     const unsigned long long* __restrict__ surv, int cap,
                                                     const int2* __restrict__ items,
                                                     const unsigned int* __restrict__ n_items,
                                                     const float* __restrict__ bound,
                                                     unsigned long long* __restrict__ kept) {
    const int lane = threadIdx.x & 63;
    const int xcd = (int)blockIdx.x % kNumXcd;
    const int nw = (int)gridDim.x / kNumXcd * 4;  // wavefronts per XCD
    const int64_t nall = (int64_t)*n_items;
    const int ni = (int)(nall * (xcd + 1) / kNumXcd);  // end of this XCD's items
    int it = (int)(nall * xcd / kNumXcd) + (int)blockIdx.x / kNumXcd * 4 +
             __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (it >= ni) return;
    // an item: (query, first pair | pairs << 16). The next item's descriptor, positions and bound
    // are loaded while the current one is chained.
    int2 item = items[it];
    int q = __builtin_amdgcn_readfirstlane(item.x);
    int j0 = __builtin_amdgcn_readfirstlane(item.y & 0xffff), c = __builtin_amdgcn_readfirstlane(item.y >> 16);
    uint32_t pos = lane < c ? (uint32_t)(surv[(int64_t)q * cap + j0 + lane] & 0xffffffffu) : 0u;
    float b = bound ? bound[q] : 3.402823466e38f;
    const int last = plan.n - 1;
    float cur[kChunkDims], nxt[kChunkDims];
    load_chunk<V4>(rows_by_pos + (int64_t)pos * d, plan.lo[0], d, cur);
    while (true) {
        const bool valid = lane < c;
        const uint32_t id = ids[pos];
        const float* __restrict__ xv = rows_by_pos + (int64_t)pos * d;
        const float* __restrict__ qv = Qp + (int64_t)q * d;
        const int itn = it + nw;
        int qn = q, j0n = 0, cn = 0;
        uint32_t posn = 0u;
        float bn = 0.f;
        if (itn < ni) {
            const int2 nitem = items[itn];
            qn = __builtin_amdgcn_readfirstlane(nitem.x);
            j0n = __builtin_amdgcn_readfirstlane(nitem.y & 0xffff);
            cn = __builtin_amdgcn_readfirstlane(nitem.y >> 16);
            posn = lane < cn ? (uint32_t)(surv[(int64_t)qn * cap + j0n + lane] & 0xffffffffu) : 0u;
            bn = bound ? bound[qn] : 3.402823466e38f;
        }
        float dis = 0.f, p = 0.f;
        // the next chunk's loads are in flight during each chain (the last chunk is peeled: its
        // chain overlaps the next item's first chunk)
        for (int ch = 0; ch < last; ++ch) {
            load_chunk<V4>(xv, plan.lo[ch + 1], d, nxt);
            chain_chunk<V4>(qv, plan.lo[ch], plan.len[ch], d, cur, p);
            if (plan.ends_block[ch]) {
                dis += p;
                p = 0.f;
            }
#pragma unroll
            for (int t = 0; t < kChunkDims; ++t) cur[t] = nxt[t];
        }
        // the next item's first chunk is in flight during the last chain of this one
        if (itn < ni) load_chunk<V4>(rows_by_pos + (int64_t)posn * d, plan.lo[0], d, nxt);
        chain_chunk<V4>(qv, plan.lo[last], plan.len[last], d, cur, p);
        dis += p;  // the last chunk closes the last block
        if (valid) kept[(int64_t)q * cap + j0 + lane] = dis <= b ? hv_key(dis, id) : HV_KEY_PAD;
        if (itn >= ni) break;
        it = itn;
        q = qn;
        j0 = j0n;
        c = cn;
        pos = posn;
        b = bn;
#pragma unroll
        for (int t = 0; t < kChunkDims; ++t) cur[t] = nxt[t];
    }
}

// The same re-rank with the rows staged through LDS (every chunk start a multiple of 4 dims, d a
// multiple of 4). Each chunk of an item's 64 rows is loaded with eight lanes per row (up to 128
// bytes of one row per eight lanes, so a load instruction touches 8 rows rather than 64: the
// texture path's cost follows the rows an instruction touches), written to this wavefront's tile
// (row stride kTileStride floats: conflict-free 16-byte row reads and chunk writes), and each lane
// then chains its own row from the tile in dimension order, as k_rerank_flat does (the same
// kept keys). The next chunk, or the next item's first, is in flight in registers during each
// chain; the rows' positions come from the lanes that hold them (cross-lane reads, no loads).
// Registers capped for 4 wavefronts per SIMD (the tiles allow 4 workgroups per CU).
constexpr int kTileStride = kChunkDims + 4;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_rerank_tile(
    const float* __restrict__ Qp, int d, const float* __restrict__ rows_by_pos, const uint32_t* __restrict__ ids,
    ChunkPlan plan, const unsigned long long* __restrict__ surv, int cap, const int2* __restrict__ items,
    const unsigned int* __restrict__ n_items, const float* __restrict__ bound, unsigned long long* __restrict__ kept) {
    __shared__ __align__(16) float s_tile[4][64 * kTileStride];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* __restrict__ tile = s_tile[w];
    const int xcd = (int)blockIdx.x % kNumXcd;
    const int nw = (int)gridDim.x / kNumXcd * 4;  // wavefronts per XCD
    const int64_t nall = (int64_t)*n_items;
    const int ni = (int)(nall * (xcd + 1) / kNumXcd);  // end of this XCD's items
    int it = (int)(nall * xcd / kNumXcd) + (int)blockIdx.x / kNumXcd * 4 + w;
    if (it >= ni) return;
    // this lane loads dims f4 .. f4 + 3 of every chunk of rows rsub + 8 t, t < 8
    const int rsub = lane >> 3, f4 = (lane & 7) * 4;
    // an item: (query, first pair | pairs << 16)
    int2 item = items[it];
    int q = __builtin_amdgcn_readfirstlane(item.x);
    int j0 = __builtin_amdgcn_readfirstlane(item.y & 0xffff), c = __builtin_amdgcn_readfirstlane(item.y >> 16);
    uint32_t pos = lane < c ? (uint32_t)(surv[(int64_t)q * cap + j0 + lane] & 0xffffffffu) : 0u;
    float b = bound ? bound[q] : 3.402823466e38f;
    const int last = plan.n - 1;
    // the rows this lane loads (rows rsub + 8 t of the item), for this item and the next
    uint32_t rp[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) rp[t] = (uint32_t)__shfl((int)pos, rsub + 8 * t, 64);
    float x[32];  // plain floats (a float4 array carried around the loop stayed in scratch)
    {
        const int col = min(plan.lo[0] + f4, d - 4);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float4 v = *reinterpret_cast<const float4*>(rows_by_pos + (int64_t)rp[t] * d + col);
            x[4 * t] = v.x;
            x[4 * t + 1] = v.y;
            x[4 * t + 2] = v.z;
            x[4 * t + 3] = v.w;
        }
    }
    while (true) {
        const bool valid = lane < c;
        const uint32_t id = ids[pos];
        const float* __restrict__ qv = Qp + (int64_t)q * d;
        const int itn = it + nw;
        int qn = q, j0n = 0, cn = 0;
        uint32_t posn = 0u;
        float bn = 0.f;
        if (itn < ni) {
            const int2 nitem = items[itn];
            qn = __builtin_amdgcn_readfirstlane(nitem.x);
            j0n = __builtin_amdgcn_readfirstlane(nitem.y & 0xffff);
            cn = __builtin_amdgcn_readfirstlane(nitem.y >> 16);
            posn = lane < cn ? (uint32_t)(surv[(int64_t)qn * cap + j0n + lane] & 0xffffffffu) : 0u;
            bn = bound ? bound[qn] : 3.402823466e38f;
        }
        uint32_t rpn[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) rpn[t] = (uint32_t)__shfl((int)posn, rsub + 8 * t, 64);
        float dis = 0.f, p = 0.f;
        for (int ch = 0; ch <= last; ++ch) {
            // the chunk's rows into the tile (every lane's reads of the previous chunk are done)
            lds_fence();
#pragma unroll
            for (int t = 0; t < 8; ++t)
                *reinterpret_cast<float4*>(tile + (rsub + 8 * t) * kTileStride + f4) =
                    make_float4(x[4 * t], x[4 * t + 1], x[4 * t + 2], x[4 * t + 3]);
            lds_fence();
            float cur[kChunkDims];
#pragma unroll
            for (int t = 0; t < kChunkDims / 4; ++t) {
                const float4 v = *reinterpret_cast<const float4*>(tile + lane * kTileStride + 4 * t);
                cur[4 * t] = v.x;
                cur[4 * t + 1] = v.y;
                cur[4 * t + 2] = v.z;
                cur[4 * t + 3] = v.w;
            }
            // the next chunk of this item, or the next item's first (row 0 when there is none), in
            // flight during this chain
            const bool nxt = ch == last;
            const int col = min((nxt ? plan.lo[0] : plan.lo[ch + 1]) + f4, d - 4);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float4 v =
                    *reinterpret_cast<const float4*>(rows_by_pos + (int64_t)(nxt ? rpn[t] : rp[t]) * d + col);
                x[4 * t] = v.x;
                x[4 * t + 1] = v.y;
                x[4 * t + 2] = v.z;
                x[4 * t + 3] = v.w;
            }
            chain_chunk<true>(qv, plan.lo[ch], plan.len[ch], d, cur, p);
            if (plan.ends_block[ch] || ch == last) {  // the last chunk closes the last block
                dis += p;
                p = 0.f;
            }
        }
        if (valid) kept[(int64_t)q * cap + j0 + lane] = dis <= b ? hv_key(dis, id) : HV_KEY_PAD;
        if (itn >= ni) break;
        it = itn;
        q = qn;
        j0 = j0n;
        c = cn;
        pos = posn;
        b = bn;
#pragma unroll
        for (int t = 0; t < 8; ++t) rp[t] = rpn[t];
    }
}

// Final top-k of a re-ranked buffer (kept keys in place, padding for dropped pairs): one
// wavefront per query compacts the kept keys into LDS (ballot prefix), sorts them in registers and
// emits the first k. Buffers of up to kSelLds pairs are loaded at once; longer ones are read in
// 64-key steps, and a query whose kept keys do not fit kSelLds is left to the workgroup select
// (done[q] = 0).
constexpr int kSelLds = 1024;  // LDS keys per wavefront: 1024, or 512 when k <= 256 (more resident)

// sort the n (<= SEL) keys of s_keys (padded), emit the first k
template <int SEL>
__device__ __forceinline__ void sort_emit_lds(const unsigned long long* s_keys, int n, int k, float* __restrict__ Dq,
                                              int64_t* __restrict__ Iq) {
    const int m = max(n, k);
    auto key_of = [&](int j) { return j < n ? s_keys[j] : HV_KEY_PAD; };
    auto emit = [&](int r, unsigned long long key) {
        if (r < k) {
            Dq[r] = hv_key_dist(key);
            Iq[r] = hv_key_id(key);
        }
    };
    if (m <= 64)
        wave_sort_emit_e<1>(m, key_of, emit);
    else if (m <= 128)
        wave_sort_emit_e<2>(m, key_of, emit);
    else if (m <= 256)
        wave_sort_emit_e<4>(m, key_of, emit);
    else if (m <= 512 || SEL <= 512)
        wave_sort_emit_e<8>(m, key_of, emit);
    else
        wave_sort_emit_e<SEL / 64>(m, key_of, emit);
}

template <int E, int SEL>
__device__ __forceinline__ void select_kept_e(const unsigned long long* __restrict__ buf, int c, int k,
                                              unsigned long long* s_keys, float* __restrict__ Dq,
                                              int64_t* __restrict__ Iq) {
    static_assert(64 * E <= SEL, "kept keys must fit the LDS row");
    const int lane = threadIdx.x & 63;
    unsigned long long v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;
        v[e] = j < c ? buf[j] : HV_KEY_PAD;
    }
    int m = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const bool keep = v[e] != HV_KEY_PAD;
        const unsigned long long b = __ballot(keep);
        if (keep) s_keys[m + __builtin_popcountll(b & ((1ull << lane) - 1ull))] = v[e];
        m += (int)__builtin_popcountll(b);
    }
    __threadfence_block();  // the keys other lanes stored are read below
    sort_emit_lds<SEL>(s_keys, m, k, Dq, Iq);
}

// SEL (512 or 1024) kept keys per wavefront fit its LDS row; the launcher takes 512 when k <= 256
// (a quarter of the LDS per workgroup, so more of them resident).
template <int SEL>
__global__ __launch_bounds__(256) void k_select_kept(const unsigned long long* __restrict__ kept,
                                                     const unsigned int* __restrict__ cnt, int64_t nq, int cap, int k,
                                                     float* __restrict__ D, int64_t* __restrict__ I,
                                                     unsigned int* __restrict__ done) {
    __shared__ unsigned long long s_keys[4][SEL];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= nq) return;
    const unsigned int cu = cnt[q];
    if (cu > (unsigned int)cap) return;  // overflowed: the exact path redoes the query
    const int c = (int)cu;
    const unsigned long long* __restrict__ buf = kept + q * cap;
    float* Dq = D + q * k;
    int64_t* Iq = I + q * k;
    bool direct = true;
    if (c <= 64)
        select_kept_e<1, SEL>(buf, c, k, s_keys[w], Dq, Iq);
    else if (c <= 128)
        select_kept_e<2, SEL>(buf, c, k, s_keys[w], Dq, Iq);
    else if (c <= 256)
        select_kept_e<4, SEL>(buf, c, k, s_keys[w], Dq, Iq);
    else if (c <= 512)
        select_kept_e<8, SEL>(buf, c, k, s_keys[w], Dq, Iq);
    else
        direct = false;
    if constexpr (SEL >= 1024) {
        if (!direct && c <= 1024) {
            select_kept_e<16, SEL>(buf, c, k, s_keys[w], Dq, Iq);
            direct = true;
        }
    }
    if (!direct) {
        int m = 0;
        for (int j0 = 0; j0 < c; j0 += 64) {
            const int j = j0 + lane;
            const unsigned long long key = j < c ? buf[j] : HV_KEY_PAD;
            const bool keep = key != HV_KEY_PAD;
            const unsigned long long b = __ballot(keep);
            const int slot = m + (int)__builtin_popcountll(b & ((1ull << lane) - 1ull));
            if (keep && slot < SEL) s_keys[w][slot] = key;
            m += (int)__builtin_popcountll(b);
        }
        if (m > SEL || k > SEL) {
            if (lane == 0) done[q] = 0u;  // the workgroup select sorts the whole buffer
            return;
        }
        __threadfence_block();
        sort_emit_lds<SEL>(s_keys[w], m, k, Dq, Iq);
    }
    if (lane == 0) done[q] = 1u;
}

void launch_rerank_items(const unsigned int* cnt, int64_t nq, int cap, const int* pairs, const int* pair_slot,
                         const int* npairs, int64_t max_pairs, int nprobe, int2* items, unsigned int* n_items,
                         int* overflow_list, int* overflow_cnt, unsigned long long* stats, hipStream_t s) {
    const int64_t threads = pairs ? max_pairs : nq;
    if (nq <= 0 || threads <= 0) return;
    hipLaunchKernelGGL(k_rerank_items, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, cnt, nq, cap, pairs,
                       pair_slot, npairs, nprobe, items, n_items, overflow_list, overflow_cnt, stats);
}

void launch_rerank(const float* Qp, int d, const float* rows_by_pos, const uint32_t* ids, const ChunkPlan& plan,
                   const unsigned long long* surv, int64_t nq, int cap, const float* bound, const int2* items,
                   const unsigned int* n_items, unsigned long long* kept, hipStream_t s) {
    if (nq <= 0) return;
    bool v4 = d >= 4 && (d & 3) == 0;
    for (int ch = 0; ch < plan.n; ++ch) v4 = v4 && (plan.lo[ch] & 3) == 0;
    const int64_t max_items = nq * ((cap + 63) / 64);
    const int64_t wgs = std::min<int64_t>((max_items + 3) / 4, kRerankGrid);
    const dim3 grid((unsigned)((wgs + kNumXcd - 1) / kNumXcd * kNumXcd));  // every XCD gets workgroups
    if (v4) {
        const int64_t twgs = std::min<int64_t>((max_items + 3) / 4, kRerankTileGrid);
        const dim3 tgrid((unsigned)((twgs + kNumXcd - 1) / kNumXcd * kNumXcd));
        hipLaunchKernelGGL(k_rerank_tile, tgrid, dim3(256), 0, s, Qp, d, rows_by_pos, ids, plan, surv, cap, items,
                           n_items, bound, kept);
    } else
        hipLaunchKernelGGL(k_rerank_flat<false>, grid, dim3(256), 0, s, Qp, d, rows_by_pos, ids, plan, surv, cap, items,
                           n_items, bound, kept);
}

void launch_select_kept(const unsigned long long* kept, const unsigned int* cnt, int64_t nq, int cap, int k, float* D,
                        int64_t* I, unsigned int* done, hipStream_t s, const RedoArgs* redo) {
    if (nq <= 0) return;
    const dim3 grid((unsigned)((nq + 3) / 4));
    const int sel = k <= 256 ? 512 : kSelLds;
    if (sel == 512)
        hipLaunchKernelGGL(k_select_kept<512>, grid, dim3(256), 0, s, kept, cnt, nq, cap, k, D, I, done);
    else
        hipLaunchKernelGGL(k_select_kept<kSelLds>, grid, dim3(256), 0, s, kept, cnt, nq, cap, k, D, I, done);
    if (cap > sel || k > sel)
        launch_select_block(kept, cnt, nq, cap, k, D, I, done, s, redo);
    else if (redo)
        launch_redo_list(*redo, s);
}

}  // namespace hv
