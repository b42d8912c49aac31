// Coarse quantizer on the matrix cores with an exact boundary re-check.
//
// The P nearest centroids of each query must come out exactly as the FP32 direct-difference
// chain orders them (acc = fma(q-c, q-c, acc) over the dimensions in order, ties to the lower
// centroid id): the plain path's probe lists and their distances are compared bit for bit.
//
// 1. k_coarse_mfma: d~(q, c) = |q|^2 + |c|^2 - 2 q.c for every (query, centroid), the dot
//    products as split BF16 products on v_mfma_f32_16x16x32_bf16 (64 x 64 tiles, 32-dim LDS
//    chunks).
// 2. k_coarse_select_wave (one wavefront per query; k_coarse_select, one workgroup per query,
//    for wide rows or large P): with err = kappa (|q|^2 + |c|^2) bounding
//    |d~ - chain| (host: kappa from d), the P-th smallest upper bound tau = P-th of (d~ + err)
//    bounds the P-th chain distance, so the P nearest all have d~ - err <= tau. Only those
//    candidates get the exact chain; their (distance, id) keys are sorted and the first P kept.
//    A query with more candidates than the buffer holds is listed for k_coarse_exact_row (all
//    centroids, exact), which strides over the list on the device.
#include <hip/hip_runtime.h>
#include <cmath>
#include <stdint.h>

#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
constexpr int kCoarseCand = 1024;  // exact re-checks per query in k_coarse_select

// Bits of the upper bound d~ + kappa (|q|^2 + |c|^2) of a chain distance (one FMA, so every
// kernel forms the same bits); negative bounds (and -0) order as 0.
__device__ __forceinline__ unsigned int coarse_ub_bits(float approx, float qq, float cc, float kappa) {
    const unsigned int u = __float_as_uint(__builtin_fmaf(kappa, qq + cc, approx));
    return (int)u < 0 ? 0u : u;
}

// The F (<= 64) smallest of the 64 E values a wavefront holds E per lane (pad 0xffffffff),
// ascending, as floats into out[0, kNearest) (lanes past F repeat the F-th): a bracket (lo, hi]
// around the F-th smallest (wave_kth_bracket), then one gather into LDS (s_tmp, 128 words of this
// wavefront) and one register sort of 128. When the bracket holds at most 64 values, the values
// <= hi are gathered (fewer than F below it, so at most 127); otherwise every value of the bracket
// equals hi (or F exceeds the values other than the padding), and the values <= lo are gathered
// with the rest of the 128 set to hi. Returns the lane's value of that order. All 64 lanes must
// call it.
template <int E>
__device__ __forceinline__ unsigned int sorted_smallest_u32(const unsigned int (&u)[E], int F, unsigned int* s_tmp,
                                                            float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const KthBracket b = wave_kth_bracket<E>(u, F);
    const bool upto_hi = b.c_hi - b.c_lo <= 64 && b.c_hi <= 128;
    const long long lim = upto_hi ? b.hi : b.lo;
    const unsigned int pad = upto_hi ? 0xffffffffu : (unsigned int)b.hi;
    int n = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const bool in = (long long)u[e] <= lim;
        const unsigned long long m = __ballot(in);
        if (in) s_tmp[n + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = u[e];
        n += __builtin_popcountll(m);
    }
    __threadfence_block();  // the words other lanes stored are read below
    unsigned int x[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) x[e] = 2 * lane + e < n ? s_tmp[2 * lane + e] : pad;
    wave_bitonic_sort<2>(x);
    // position p of the order lives in lane p / 2, slot p % 2
    const int p = min(lane, F - 1);
    const unsigned int y0 = (unsigned int)__shfl((int)x[0], p >> 1, 64);
    const unsigned int y1 = (unsigned int)__shfl((int)x[1], p >> 1, 64);
    const unsigned int val = (p & 1) ? y1 : y0;
    out[lane] = __uint_as_float(val);
    return val;
}

// squared norms of rows (any summation order: kappa covers it)
__global__ void k_row_norms(const float* __restrict__ X, int64_t n, int d, float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const int lane = threadIdx.x & 63;
    float s = 0.f;
    for (int i = lane; i < d; i += 64) {
        const float v = X[r * d + i];
        s = __builtin_fmaf(v, v, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[r] = s;
}

void launch_row_norms(const float* X, int64_t n, int d, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_row_norms, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, X, n, d, out);
}

// 64 queries x 64 centroids per workgroup; wavefront w computes queries 16w..16w+15 against the
// 64 centroids (four 16 x 16 MFMA tiles). Lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[4(l/16)+i][l%16]. The tiles pass through LDS 32 dims at a time; with d % 4 == 0 each thread
// loads 2 + 2 float4 per chunk and the next chunk's loads are in flight while the current one is
// multiplied, and the squared norms of the 64 + 64 rows are summed from the same loads (8 threads
// per row, any order: kappa covers it): the workgroups of the first column write the query norms,
// those of the first row the centroid norms, for the select. With d % 4 != 0 the norms come from
// k_row_norms, launched before. The search's counters and buffers that start cleared (fill) are
// cleared by the same grid first (a grid-stride loop; no kernel of this launch reads them), so a
// search needs no launch of its own for them.
__global__ __launch_bounds__(256) void k_coarse_mfma(const float* __restrict__ Q, int64_t nq,
                                                     const float* __restrict__ C, int nc, int d,
                                                     float* __restrict__ qn, float* __restrict__ cn,
                                                     float* __restrict__ out, FillSegs fill) {
    if (fill.nseg > 0) {
        const int64_t stride = (int64_t)gridDim.x * gridDim.y * blockDim.x;
        const int64_t g0 = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
        for (int j = 0; j < fill.nseg; ++j) {
            int* p = fill.p[j];
            const int64_t n = fill.n[j];
            const int v = fill.v[j];
            for (int64_t i = g0; i < n; i += stride) p[i] = v;
        }
    }
    // row stride 36 floats: a row starts 16-byte aligned, and the 8 rows one 16-byte LDS read
    // serves at a time start 4 banks apart (36 mod 32), so each such read is conflict-free
    constexpr int KC = 32, LS = KC + 4;
    __shared__ __align__(16) float Qs[64 * LS];
    __shared__ __align__(16) float Cs[64 * LS];
    __shared__ float s_qn[64], s_cn[64];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int64_t q0 = (int64_t)blockIdx.y * 64;
    const int c0 = blockIdx.x * 64;
    floatx4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // One chunk of 32 dims is one K-step of v_mfma_f32_16x16x32_bf16: lane l holds dims
    // 8 (l / 16) .. +8 of its row (2 x 16 bytes from LDS), split into BF16 parts v = hi + lo + e
    // (hi the BF16 rounding of v, lo that of v - hi, which is exact in FP32; |e| <= 2^-16 |v|),
    // and q.c is taken as qh.ch + qh.cl + ql.ch: three MFMAs whose products are exact in FP32
    // (coarse_kappa bounds what is left out and the FP32 accumulation)
    auto split8 = [](const float4* p, bf16x8& hi, bf16x8& lo) {
        const float4 x0 = p[0], x1 = p[1];
        const float v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            hi[e] = (__bf16)v[e];
            lo[e] = (__bf16)(v[e] - (float)hi[e]);
        }
    };
    auto mult = [&]() {
        bf16x8 ah, al;
        split8(reinterpret_cast<const float4*>(Qs + (16 * w + (l & 15)) * LS + 8 * (l >> 4)), ah, al);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bf16x8 bh, bl;
            split8(reinterpret_cast<const float4*>(Cs + (16 * j + (l & 15)) * LS + 8 * (l >> 4)), bh, bl);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[j], 0, 0, 0);
        }
    };
    if ((d & 3) == 0) {
        // float4 e = t + 256 i (i = 0, 1) of a chunk: row e / 8, dims k0 + 4 (e % 8) .. +3
        float4 qv[2], cv[2];
        float sq[2] = {0.f, 0.f}, sc[2] = {0.f, 0.f};
        auto load = [&](int k0) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e = t + 256 * i, r = e >> 3, k = k0 + 4 * (e & 7);
                qv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                cv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (k < d) {
                    if (q0 + r < nq) qv[i] = *reinterpret_cast<const float4*>(Q + (q0 + r) * d + k);
                    if (c0 + r < nc) cv[i] = *reinterpret_cast<const float4*>(C + (int64_t)(c0 + r) * d + k);
                }
            }
        };
        load(0);
        for (int k0 = 0; k0 < d; k0 += KC) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e = t + 256 * i, r = e >> 3, kk = 4 * (e & 7);
                sq[i] += qv[i].x * qv[i].x + qv[i].y * qv[i].y + qv[i].z * qv[i].z + qv[i].w * qv[i].w;
                sc[i] += cv[i].x * cv[i].x + cv[i].y * cv[i].y + cv[i].z * cv[i].z + cv[i].w * cv[i].w;
                *reinterpret_cast<float4*>(Qs + r * LS + kk) = qv[i];
                *reinterpret_cast<float4*>(Cs + r * LS + kk) = cv[i];
            }
            __syncthreads();
            if (k0 + KC < d) load(k0 + KC);  // next chunk in flight during the MFMAs
            mult();
            __syncthreads();
        }
        // row r = (t >> 3) + 32 i is held by the 8 consecutive threads 8 (t >> 3) .. +7
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int o = 1; o < 8; o <<= 1) {
                sq[i] += __shfl_xor(sq[i], o, 64);
                sc[i] += __shfl_xor(sc[i], o, 64);
            }
            const int r = (t >> 3) + 32 * i;
            if ((t & 7) == 0) {
                s_qn[r] = sq[i];
                s_cn[r] = sc[i];
                if (blockIdx.x == 0 && q0 + r < nq) qn[q0 + r] = sq[i];
                if (blockIdx.y == 0 && c0 + r < nc) cn[c0 + r] = sc[i];
            }
        }
    } else {
        for (int k0 = 0; k0 < d; k0 += KC) {
            const int kn = min(KC, d - k0);
            for (int e = t; e < 64 * KC; e += 256) {
                const int r = e / KC, kk = e % KC;
                float v = 0.f, u = 0.f;
                if (kk < kn) {
                    if (q0 + r < nq) v = Q[(q0 + r) * d + k0 + kk];
                    if (c0 + r < nc) u = C[(int64_t)(c0 + r) * d + k0 + kk];
                }
                Qs[r * LS + kk] = v;
                Cs[r * LS + kk] = u;
            }
            __syncthreads();
            mult();
            __syncthreads();
        }
        if (t < 64) {
            s_qn[t] = q0 + t < nq ? qn[q0 + t] : 0.f;
            s_cn[t] = c0 + t < nc ? cn[c0 + t] : 0.f;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = c0 + 16 * j + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t q = q0 + 16 * w + 4 * (l >> 4) + i;
            if (q < nq && c < nc) out[q * nc + c] = s_qn[q - q0] + s_cn[16 * j + (l & 15)] - 2.f * acc[j][i];
        }
    }
}

// The exact chain of one (query, centroid) pair, acc = fma(q-c, q-c, acc) over the dimensions in
// order. The centroid row is loaded 16 dims at a time (16-byte loads when d % 4 == 0), the next
// 16 in flight while the current ones are chained (32 registers for both: the select kernels
// keep their occupancy); q is read where it lies (LDS, or memory with a wavefront-uniform
// address).
constexpr int kChainDims = 16;

__device__ __forceinline__ void coarse_load(const float* __restrict__ row, int lo, int d, bool v4,
                                            float (&r)[kChainDims]) {
    if (v4) {
#pragma unroll
        for (int t = 0; t < kChainDims / 4; ++t) {
            const float4 x = *reinterpret_cast<const float4*>(row + min(lo + 4 * t, d - 4));
            r[4 * t] = x.x;
            r[4 * t + 1] = x.y;
            r[4 * t + 2] = x.z;
            r[4 * t + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < kChainDims; ++t) r[t] = row[min(lo + t, d - 1)];
    }
}

__device__ __forceinline__ float coarse_chain(const float* __restrict__ qs, const float* __restrict__ crow, int d) {
    const bool v4 = (d & 3) == 0;
    const int last = (d - 1) / kChainDims;
    float cur[kChainDims], nxt[kChainDims];
    coarse_load(crow, 0, d, v4, cur);
    float acc = 0.f;
    for (int ch = 0; ch < last; ++ch) {
        coarse_load(crow, (ch + 1) * kChainDims, d, v4, nxt);
#pragma unroll
        for (int t = 0; t < kChainDims; ++t) {
            const float u = qs[ch * kChainDims + t] - cur[t];
            acc = __builtin_fmaf(u, u, acc);
        }
#pragma unroll
        for (int t = 0; t < kChainDims; ++t) cur[t] = nxt[t];
    }
    const int lo = last * kChainDims, len = d - lo;
#pragma unroll
    for (int t = 0; t < kChainDims; ++t) {
        if (t < len) {
            const float u = qs[lo + t] - cur[t];
            acc = __builtin_fmaf(u, u, acc);
        }
    }
    return acc;
}

constexpr int kSelPer = kMaxSelectRow / 256;  // centroids per thread (nlist <= 8192)

__global__ __launch_bounds__(256) void k_coarse_select(const float* __restrict__ approx, const float* __restrict__ Q,
                                                       const float* __restrict__ C, int nc, int d,
                                                       const float* __restrict__ qn, const float* __restrict__ cn,
                                                       float kappa, int P, int* __restrict__ probes,
                                                       float* __restrict__ pdist, int* __restrict__ redo_list,
                                                       int* __restrict__ redo_cnt) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(s_raw);          // [kCoarseCand]
    float* qs = reinterpret_cast<float*>(s_raw + sizeof(unsigned long long) * kCoarseCand);  // [d]
    __shared__ unsigned int s_hist[256];
    __shared__ unsigned int s_sel[2];
    __shared__ unsigned int s_red[2][4];
    __shared__ int s_cnt;
    const int64_t q = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const float* __restrict__ row = approx + q * nc;
    const float qq = qn[q];
    // upper bounds (FP32 bits of non-negative values order as unsigned ints)
    unsigned int ub[kSelPer];
    float lbv[kSelPer];
    unsigned int lo = 0xffffffffu, hi = 0u;
#pragma unroll
    for (int i = 0; i < kSelPer; ++i) {
        const int c = t + 256 * i;
        ub[i] = 0xffffffffu;
        lbv[i] = 3.402823466e38f;
        if (c < nc) {
            const float dt = row[c];
            const float err = kappa * (qq + cn[c]);
            const float u = fmaxf(dt + err, 0.f);
            lbv[i] = dt - err;
            ub[i] = __float_as_uint(u);
            if (ub[i] == 0x80000000u) ub[i] = 0u;
            lo = min(lo, ub[i]);
            hi = max(hi, ub[i]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (unsigned int)__shfl_xor((int)lo, o, 64));
        hi = max(hi, (unsigned int)__shfl_xor((int)hi, o, 64));
    }
    if (lane == 0) {
        s_red[0][w] = lo;
        s_red[1][w] = hi;
    }
    for (int i = t; i < d; i += 256) qs[i] = Q[q * d + i];
    __syncthreads();
    lo = min(min(s_red[0][0], s_red[0][1]), min(s_red[0][2], s_red[0][3]));
    hi = max(max(s_red[1][0], s_red[1][1]), max(s_red[1][2], s_red[1][3]));
    // tau = P-th smallest upper bound (radix select from the first byte where lo and hi differ)
    const unsigned int diff = lo ^ hi;
    const int top = diff ? (31 - __builtin_clz(diff)) / 8 * 8 : -8;
    unsigned int mask = top >= 24 ? 0u : (0xffffffffu << (top + 8));
    unsigned int prefix = lo & mask;
    int kk = P;
    for (int shift = top; shift >= 0; shift -= 8) {
        s_hist[t] = 0u;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSelPer; ++i)
            if (t + 256 * i < nc && (ub[i] & mask) == prefix) atomicAdd(&s_hist[(ub[i] >> shift) & 255u], 1u);
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
    const float tau = __uint_as_float(prefix);
    // the candidates: lower bound <= tau; their exact chain distances as (distance, id) keys
    if (t == 0) s_cnt = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSelPer; ++i) {
        const int c = t + 256 * i;
        if (c < nc && lbv[i] <= tau) {
            const int pos = atomicAdd(&s_cnt, 1);
            if (pos < kCoarseCand) keys[pos] = (unsigned long long)(unsigned int)c;
        }
    }
    __syncthreads();
    const int m = s_cnt;
    if (m > kCoarseCand) {  // too wide a band: the exact kernel redoes this query
        if (t == 0) redo_list[atomicAdd(redo_cnt, 1)] = (int)q;
        return;
    }
    for (int e = t; e < m; e += 256) {
        const int c = (int)keys[e];
        keys[e] = hv_key(coarse_chain(qs, C + (int64_t)c * d, d), (uint32_t)c);
    }
    const int p2 = hv_next_pow2(max(m, 2));
    for (int e = m + t; e < p2; e += 256) keys[e] = HV_KEY_PAD;
    __syncthreads();
    block_bitonic_sort(keys, p2);
    for (int r = t; r < P; r += 256) {
        probes[q * P + r] = (int)hv_key_id(keys[r]);
        pdist[q * P + r] = hv_key_dist(keys[r]);
    }
}

// The same selection with one wavefront per query (P <= kWaveP, nc <= 64 * E): the upper bounds
// live in registers, tau comes from wave_kth_u32, the candidates' exact keys (at most CAND, else
// the query is listed for k_coarse_exact_row) are sorted in registers. CAND = 256 for P <= 128
// (less LDS, more resident wavefronts), 512 beyond.
constexpr int kWaveP = 256;

template <int EC>
__device__ __forceinline__ void coarse_finish_wave(const unsigned long long* __restrict__ keys, int m, int P,
                                                   int* __restrict__ probes, float* __restrict__ pdist,
                                                   int* __restrict__ bucket_cnt, int nranks) {
    const int lane = threadIdx.x & 63;
    unsigned long long v[EC];
#pragma unroll
    for (int e = 0; e < EC; ++e) {
        const int j = e * 64 + lane;
        v[e] = j < m ? keys[j] : HV_KEY_PAD;
    }
    wave_bitonic_sort<EC>(v);
#pragma unroll
    for (int e = 0; e < EC; ++e) {
        const int r = lane * EC + e;
        if (r < P) {
            const int id = (int)hv_key_id(v[e]);
            probes[r] = id;
            pdist[r] = hv_key_dist(v[e]);
            if (bucket_cnt && r < nranks) atomicAdd(&bucket_cnt[(int64_t)id * nranks + r], 1);
        }
    }
}

template <int E, int CAND>
__global__ __launch_bounds__(256) void k_coarse_select_wave(const float* __restrict__ approx,
                                                            const float* __restrict__ Q,
                                                            const float* __restrict__ C, int nc, int d,
                                                            const float* __restrict__ qn,
                                                            const float* __restrict__ cn, float kappa, int64_t nq,
                                                            int P, int* __restrict__ probes,
                                                            float* __restrict__ pdist, int* __restrict__ redo_list,
                                                            int* __restrict__ redo_cnt, float* __restrict__ near,
                                                            int* __restrict__ bucket_cnt, int nranks) {
    __shared__ unsigned int s_cand[4][CAND];
    __shared__ unsigned long long s_ckey[4][CAND];
    __shared__ unsigned int s_tmp[4][128];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= nq) return;
    const float* __restrict__ row = approx + q * nc;
    const float qq = qn[q];
    unsigned int ub[E];
    float lbf[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int c = e * 64 + lane;
        const float a = c < nc ? row[c] : 0.f, cc = c < nc ? cn[c] : 0.f;
        ub[e] = c < nc ? coarse_ub_bits(a, qq, cc, kappa) : 0xffffffffu;
        lbf[e] = c < nc ? a - kappa * (qq + cc) : __builtin_inff();
    }
    float tau;
    if (near) {
        // the kNearest smallest upper bounds, sorted (the learned features; P <= kNearest): their
        // P-th is tau
        const unsigned int v = sorted_smallest_u32<E>(ub, min(nc, kNearest), s_tmp[w], near + q * kNearest);
        tau = __uint_as_float((unsigned int)__shfl((int)v, P - 1, 64));
    } else {
        tau = __uint_as_float(wave_kth_u32<E>(ub, P, s_tmp[w]));
    }
    // the candidates: lower bound <= tau (the bounds were kept in registers through the search:
    // re-reading the row costs more than the occupancy they take)
    int m = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int c = e * 64 + lane;
        const bool in = c < nc && lbf[e] <= tau;
        const unsigned long long b = __ballot(in);
        const int slot = m + (int)__builtin_popcountll(b & ((1ull << lane) - 1ull));
        if (in && slot < CAND) s_cand[w][slot] = (unsigned int)c;
        m += (int)__builtin_popcountll(b);
    }
    if (m > CAND) {  // too wide a band: k_coarse_exact_row redoes this query (and counts its pairs)
        if (lane == 0) redo_list[atomicAdd(redo_cnt, 1)] = (int)q;
        return;
    }
    const float* __restrict__ qv = Q + q * d;
    int* pq = probes + q * P;
    float* dq = pdist + q * P;
    __threadfence_block();  // the candidates other lanes stored are read below
    // exact keys of the candidates, one chain at a time per lane
#pragma unroll 1
    for (int j = lane; j < m; j += 64) {
        const unsigned int c = s_cand[w][j];
        s_ckey[w][j] = hv_key(coarse_chain(qv, C + (int64_t)c * d, d), c);
    }
    __threadfence_block();
    if (m <= 64)
        coarse_finish_wave<1>(s_ckey[w], m, P, pq, dq, bucket_cnt, nranks);
    else if (m <= 128)
        coarse_finish_wave<2>(s_ckey[w], m, P, pq, dq, bucket_cnt, nranks);
    else if (CAND <= 256 || m <= 256)
        coarse_finish_wave<4>(s_ckey[w], m, P, pq, dq, bucket_cnt, nranks);
    else
        coarse_finish_wave<8>(s_ckey[w], m, P, pq, dq, bucket_cnt, nranks);
}

// Fallback: every centroid's exact chain, LDS sort, first P (one workgroup per listed query).
__global__ __launch_bounds__(256) void k_coarse_exact_row(const int* __restrict__ redo_list,
                                                          const int* __restrict__ redo_cnt,
                                                          const float* __restrict__ Q, const float* __restrict__ C,
                                                          int nc, int d, int P, int* __restrict__ probes,
                                                          float* __restrict__ pdist, int* __restrict__ bucket_cnt,
                                                          int nranks) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    const int p2 = hv_next_pow2(max(nc, 2));
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(s_raw);  // [p2]
    float* qs = reinterpret_cast<float*>(s_raw + sizeof(unsigned long long) * p2);
    const int cnt = *redo_cnt;
    for (int w = blockIdx.x; w < cnt; w += gridDim.x) {
        const int64_t q = redo_list[w];
        __syncthreads();
        for (int i = threadIdx.x; i < d; i += blockDim.x) qs[i] = Q[q * d + i];
        __syncthreads();
        for (int c = threadIdx.x; c < p2; c += blockDim.x)
            keys[c] = c < nc ? hv_key(coarse_chain(qs, C + (int64_t)c * d, d), (uint32_t)c) : HV_KEY_PAD;
        __syncthreads();
        block_bitonic_sort(keys, p2);
        for (int r = threadIdx.x; r < P; r += blockDim.x) {
            const int id = (int)hv_key_id(keys[r]);
            probes[q * P + r] = id;
            pdist[q * P + r] = hv_key_dist(keys[r]);
            if (bucket_cnt && r < nranks) atomicAdd(&bucket_cnt[(int64_t)id * nranks + r], 1);
        }
    }
}

// The kNearest smallest distance values of each row of v [n][nc], ascending (a row of fewer
// columns repeats its largest): the learned predictor's distance features. With qn (coarse
// estimate rows from k_coarse_mfma), the values are the upper bounds d~ + kappa (|q|^2 + |c|^2)
// that k_coarse_select_wave also forms (the same bits, so features do not depend on which kernel
// produced them); without, the rows are exact distances, negatives as 0. One wavefront per row
// (nc <= 64 E).
template <int E>
__global__ __launch_bounds__(256) void k_near_rows_wave(const float* __restrict__ v, int64_t n, int nc,
                                                        const float* __restrict__ qn, const float* __restrict__ cn,
                                                        float kappa, float* __restrict__ out) {
    __shared__ unsigned int s_tmp[4][128];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + w;
    if (r >= n) return;
    const float* __restrict__ row = v + r * nc;
    const float qq = qn ? qn[r] : 0.f;
    unsigned int u[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int c = e * 64 + lane;
        if (c >= nc) {
            u[e] = 0xffffffffu;
        } else if (qn) {
            u[e] = coarse_ub_bits(row[c], qq, cn[c], kappa);
        } else {
            const float x = row[c];
            u[e] = x > 0.f ? __float_as_uint(x) : 0u;  // non-negative floats order as uints
        }
    }
    sorted_smallest_u32<E>(u, min(nc, kNearest), s_tmp[w], out + r * kNearest);
}

// The same for rows of up to kMaxSelectRow columns: one workgroup per row sorts (value, column)
// keys in LDS.
__global__ __launch_bounds__(256) void k_near_rows_block(const float* __restrict__ v, int64_t n, int nc,
                                                         const float* __restrict__ qn, const float* __restrict__ cn,
                                                         float kappa, float* __restrict__ out) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t r = blockIdx.x;
    const int M = hv_next_pow2(max(nc, 2));
    for (int i = threadIdx.x; i < M; i += blockDim.x) {
        unsigned int u = 0u;
        if (i < nc) {
            const float x = v[r * nc + i];
            u = qn ? coarse_ub_bits(x, qn[r], cn[i], kappa) : (x > 0.f ? __float_as_uint(x) : 0u);
        }
        s_keys[i] = i < nc ? (((unsigned long long)u << 32) | (unsigned int)i) : ~0ull;
    }
    block_bitonic_sort(s_keys, M);
    const int F = min(nc, kNearest);
    for (int i = threadIdx.x; i < kNearest; i += blockDim.x)
        out[r * kNearest + i] = __uint_as_float((unsigned int)(s_keys[min(i, F - 1)] >> 32));
}

void launch_near_rows(const float* v, int64_t n, int nc, const float* qn, const float* cn, float kappa, float* out,
                      hipStream_t s) {
    if (n <= 0 || nc <= 0) return;
    const dim3 wgrid((unsigned)((n + 3) / 4));
#define HV_NEAR(E) hipLaunchKernelGGL(k_near_rows_wave<E>, wgrid, dim3(256), 0, s, v, n, nc, qn, cn, kappa, out)
    if (nc <= 512)
        HV_NEAR(8);
    else if (nc <= 1024)
        HV_NEAR(16);
    else if (nc <= 2048)
        HV_NEAR(32);
    else if (nc <= 4096)
        HV_NEAR(64);
    else
        hipLaunchKernelGGL(k_near_rows_block, dim3((unsigned)n), dim3(256),
                           sizeof(unsigned long long) * hv_next_pow2(nc), s, v, n, nc, qn, cn, kappa, out);
#undef HV_NEAR
}

float coarse_kappa(int d) {
    // |d~ - chain| <= kappa (|q|^2 + |c|^2), every term taken against |q||c| <= (|q|^2 + |c|^2) / 2
    // where it multiplies |q||c|:
    // - the split dot product (BF16 parts: |ql| <= 2^-8 |q|, |eq| <= 2^-16 |q|, and so for c):
    //   3d exact products accumulated in FP32 in the MFMA's order (gamma_6d, twice gamma_3d for
    //   the unknown order) over their absolute sum (<= 1.016 |q||c|), and the parts left out,
    //   ql.cl + (qh + ql).ec + eq.c (<= 3.01 2^-16 |q||c|), both doubled by the factor 2 of the
    //   dot product: 6.1 d u + 3.1 2^-16;
    // - the FP32 norms (gamma_d each): d u;
    // - the chain's own error against the real distance (2 gamma_2d of 2 (|q|^2 + |c|^2)): 8 d u;
    // - the three roundings of d~ and slack: (d + 8) u;
    // and 2 % for the rounding of this bound.
    const double u = 1.0 / 16777216.0;
    return (float)(1.02 * ((16.2 * d + 8.0) * u + 3.1 * std::ldexp(1.0, -16)));
}

void launch_coarse_mfma(const float* Q, int64_t nq, const float* C, int nc, int d, float* qn, float* cn,
                        float* approx, hipStream_t s, const FillSegs* fill) {
    FillSegs f{};
    if (fill) f = *fill;
    if (nq <= 0) {
        launch_fill_multi(f, s);
        return;
    }
    if ((d & 3) != 0) {  // the kernel sums the norms itself only from its 16-byte loads
        launch_row_norms(C, nc, d, cn, s);
        launch_row_norms(Q, nq, d, qn, s);
    }
    const dim3 grid((unsigned)((nc + 63) / 64), (unsigned)((nq + 63) / 64));
    hipLaunchKernelGGL(k_coarse_mfma, grid, dim3(256), 0, s, Q, nq, C, nc, d, qn, cn, approx, f);
}

bool launch_coarse_select(const float* approx, const float* Q, const float* C, int nc, int d, const float* qn,
                          const float* cn, float kappa, int64_t nq, int P, int* probes, float* pdist,
                          int* redo_list, int* redo_cnt, float* near, int* bucket_cnt, int nranks, hipStream_t s) {
    if (nq <= 0) return bucket_cnt != nullptr;
    const dim3 wgrid((unsigned)((nq + 3) / 4));
    // the wavefront kernel forms the features itself when P <= kNearest
    float* fused = P <= kNearest ? near : nullptr;
#define HV_CSEL(E, CAND)                                                                                        \
    hipLaunchKernelGGL((k_coarse_select_wave<E, CAND>), wgrid, dim3(256), 0, s, approx, Q, C, nc, d, qn, cn, kappa, \
                       nq, P, probes, pdist, redo_list, redo_cnt, fused, bucket_cnt, nranks)
    const bool narrow = P <= 128;
    if (P <= kWaveP && nc <= 512) {
        if (narrow) HV_CSEL(8, 256); else HV_CSEL(8, 512);
    } else if (P <= kWaveP && nc <= 1024) {
        if (narrow) HV_CSEL(16, 256); else HV_CSEL(16, 512);
    } else if (P <= kWaveP && nc <= 2048) {
        if (narrow) HV_CSEL(32, 256); else HV_CSEL(32, 512);
    } else {
        fused = nullptr;
        bucket_cnt = nullptr;  // the caller counts the pairs
        const size_t lds = sizeof(unsigned long long) * kCoarseCand + sizeof(float) * ((d + 3) & ~3);
        hipLaunchKernelGGL(k_coarse_select, dim3((unsigned)nq), dim3(256), lds, s, approx, Q, C, nc, d, qn, cn, kappa,
                           P, probes, pdist, redo_list, redo_cnt);
    }
#undef HV_CSEL
    if (near && !fused) launch_near_rows(approx, nq, nc, qn, cn, kappa, near, s);
    // the queries whose band overflowed its candidate buffer, exactly
    const size_t lds2 = sizeof(unsigned long long) * hv_next_pow2(nc > 2 ? nc : 2) + sizeof(float) * ((d + 3) & ~3);
    hipLaunchKernelGGL(k_coarse_exact_row, dim3((unsigned)std::min<int64_t>(nq, 256)), dim3(256), lds2, s, redo_list,
                       redo_cnt, Q, C, nc, d, P, probes, pdist, bucket_cnt, nranks);
    return bucket_cnt != nullptr;
}

}  // namespace hv
