// Device-side helpers shared by the HIP kernels of the IVF-flat search path (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// A candidate is carried as one 64-bit key: FP32 distance bits in the high word, the vector
// id in the low word. Distances are >= 0 (sums of squares), so the unsigned order of the
// bit patterns is the numeric order and the key order is the (distance, id) lexicographic
// order the results are defined by. The padding entry (FLT_MAX, -1) of the reference's
// heap_init (heap.hpp:13-25) is the key below; it sorts after every real entry.
#define HV_KEY_PAD 0x7f7fffffffffffffull
#define HV_ID_NONE 0xffffffffu

__device__ __forceinline__ uint64_t hv_key(float dist, uint32_t id) {
    return ((uint64_t)__float_as_uint(dist) << 32) | (uint64_t)id;
}
__device__ __forceinline__ float hv_key_dist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ int64_t hv_key_id(uint64_t k) {
    uint32_t lo = (uint32_t)(k & 0xffffffffu);
    return lo == HV_ID_NONE ? (int64_t)-1 : (int64_t)lo;
}

// Inner-product keys: the high word is a decreasing map of the FP32 inner product onto uint32
// (larger products first), the low word the id, so ascending keys are (inner product descending,
// id ascending); -0.0 counts as +0.0. HV_IP_KEY_PAD sorts after every real key.
#define HV_IP_KEY_PAD 0xffffffffffffffffull
__device__ __forceinline__ uint64_t hv_ip_key(float ip, uint32_t id) {
    const uint32_t b = __float_as_uint(ip == 0.f ? 0.f : ip);
    const uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);  // increasing with ip
    return ((uint64_t)(~o) << 32) | (uint64_t)id;
}
__device__ __forceinline__ float hv_ip_key_value(uint64_t k) {
    const uint32_t o = ~(uint32_t)(k >> 32);
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Workgroup-wide ascending bitonic sort of n keys in LDS (n a power of two, n >= 2).
// Every thread of the workgroup must call it. Ends with a barrier.
template <typename T>
__device__ void block_bitonic_sort(T* s, int n) {
    const int nt = blockDim.x;
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < (n >> 1); t += nt) {
                int i = 2 * t - (t & (stride - 1));
                int j = i + stride;
                bool up = (i & size) == 0;
                T a = s[i], b = s[j];
                if ((a > b) == up) {
                    s[i] = b;
                    s[j] = a;
                }
            }
        }
    }
    __syncthreads();
}

// Workgroup barrier for LDS traffic only: this wavefront's LDS operations complete, then
// s_barrier. Global loads in flight stay in flight (a __syncthreads() waits for them as well,
// which drains a software pipeline of prefetched global data at every barrier).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Orders this wavefront's LDS accesses (stores by some lanes, then loads of them by others)
// without waiting for its global loads.
__device__ __forceinline__ void lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup", "local"); }

// e / den for an index e >= 0 and 0 < den < 2^31: the 32-bit unsigned division (a short
// reciprocal-based sequence) when e < 2^32, the 64-bit one (a long software sequence) otherwise.
__device__ __forceinline__ int64_t hv_idiv(int64_t e, int den) {
    return ((uint64_t)e >> 32) == 0 ? (int64_t)((uint32_t)e / (uint32_t)den) : e / (int64_t)den;
}

// A query's own k when per-query counts are given (clamped to [1, k]), else k.
__device__ __forceinline__ int hv_query_k(const int* kq, int64_t q, int k) {
    return kq ? max(1, min(kq[q], k)) : k;
}

__host__ __device__ inline int hv_next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// The value of lane ^ m, for m a power of two known at compile time after unrolling: a DPP
// quad permutation for 1 and 2 (a VALU move), a bit-mode swizzle within 32 lanes for 4, 8 and 16
// (new lane = lane ^ m, no address operand), a permute across the halves for 32.
__device__ __forceinline__ unsigned int shfl_xor_pow2(unsigned int x, int m) {
    switch (m) {
        case 1: return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xf, 0xf, false);  // [1,0,3,2]
        case 2: return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xf, 0xf, false);  // [2,3,0,1]
        case 4: return (unsigned int)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (4 << 10));
        case 8: return (unsigned int)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (8 << 10));
        case 16: return (unsigned int)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (16 << 10));
        default: return (unsigned int)__shfl_xor((int)x, m, 64);
    }
}
__device__ __forceinline__ unsigned long long shfl_xor_pow2(unsigned long long x, int m) {
    const unsigned int lo = shfl_xor_pow2((unsigned int)x, m), hi = shfl_xor_pow2((unsigned int)(x >> 32), m);
    return (unsigned long long)hi << 32 | lo;
}

// Wavefront-wide ascending bitonic sort of 64*E keys held in registers: element i lives in
// lane i / E, slot i % E. Strides below E exchange within a lane; larger strides exchange with
// lane ^ (stride / E) through cross-lane moves (shfl_xor_pow2). No LDS, no barriers. T is
// unsigned int or unsigned long long.
template <int E, typename T>
__device__ __forceinline__ void wave_bitonic_sort(T (&v)[E]) {
    const int lane = threadIdx.x & 63;
    constexpr int N = 64 * E;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= E) {
                const int lx = stride / E;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = lane * E + e;
                    const T o = shfl_xor_pow2(v[e], lx);
                    const bool up = (i & size) == 0;
                    const bool lower = (i & stride) == 0;
                    const T mn = o < v[e] ? o : v[e];
                    const T mx = o < v[e] ? v[e] : o;
                    v[e] = (lower == up) ? mn : mx;
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & stride) continue;
                    const int e2 = e | stride;
                    const bool up = ((lane * E + e) & size) == 0;
                    const T a = v[e], b = v[e2];
                    const bool sw = (a > b) == up;
                    v[e] = sw ? b : a;
                    v[e2] = sw ? a : b;
                }
            }
        }
    }
}

// out[(*cnt)++] = key for every lane with pred, one LDS atomic per wavefront (the lanes' slots
// follow their order in the wavefront). Every lane of the wavefront that is active must call it.
__device__ __forceinline__ void wave_append(bool pred, unsigned long long key, unsigned long long* out, int* cnt) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(m);
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __builtin_popcountll(m));
    base = __shfl(base, leader, 64);
    if (pred) out[base + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = key;
}

// k-th smallest (1-based) of the na keys of a and the nb keys of b (block-wide; every thread
// gets it; na + nb < 2^24). Radix select over the bytes of the 64-bit keys, from the highest byte
// in which the smallest and the largest key differ (the bytes above it are common to every key,
// so their passes would put every key in one bin); it stops as soon as the rank-k bin holds
// exactly the keys of ranks <= k that remain (the k-th is then the largest key with that prefix,
// one max reduction), so distinct distances take two or three passes instead of eight.
__device__ inline unsigned long long block_radix_kth(const unsigned long long* a, int na, const unsigned long long* b,
                                                     int nb, int k, unsigned int* hist,
                                                     unsigned long long* shared_out) {
    const int n = na + nb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = (blockDim.x + 63) >> 6;
    unsigned long long* red = reinterpret_cast<unsigned long long*>(hist);  // 2 words per wavefront
    unsigned long long mn = ~0ull, mx = 0ull;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const unsigned long long key = i < na ? a[i] : b[i - na];
        mn = key < mn ? key : mn;
        mx = key > mx ? key : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(mn, o, 64), y = __shfl_xor(mx, o, 64);
        mn = x < mn ? x : mn;
        mx = y > mx ? y : mx;
    }
    if (lane == 0) {
        red[2 * wave] = mn;
        red[2 * wave + 1] = mx;
    }
    __syncthreads();
    mn = red[0];
    mx = red[1];
    for (int w = 1; w < nwave; ++w) {
        mn = red[2 * w] < mn ? red[2 * w] : mn;
        mx = red[2 * w + 1] > mx ? red[2 * w + 1] : mx;
    }
    __syncthreads();  // red shares hist's words
    if (mn == mx) return mn;
    int shift = ((63 - __builtin_clzll(mn ^ mx)) >> 3) << 3;
    unsigned long long mask = shift + 8 >= 64 ? 0ull : ~0ull << (shift + 8);
    unsigned long long prefix = mn & mask;
    int kk = k;
    for (; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const unsigned long long key = i < na ? a[i] : b[i - na];
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255ull], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // wavefront 0: 4 bins per lane, inclusive scan, find the bin holding rank kk
            unsigned int h[4], sum = 0u;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                h[t] = hist[lane * 4 + t];
                sum += h[t];
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
                unsigned int below = run + h[0] + h[1] + h[2], cnt = h[3];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (run + h[t] >= (unsigned int)kk) {
                        digit = lane * 4 + t;
                        below = run;
                        cnt = h[t];
                        break;
                    }
                    run += h[t];
                }
                shared_out[0] = ((unsigned long long)digit << 48) | ((unsigned long long)cnt << 24) | below;
            }
        }
        __syncthreads();
        const unsigned long long r = shared_out[0];
        kk -= (int)(r & 0xffffffull);
        prefix |= ((r >> 48) & 255ull) << shift;
        mask |= 255ull << shift;
        __syncthreads();
        if (kk == (int)((r >> 24) & 0xffffffull)) {
            // every key with this prefix has rank <= k: the k-th is the largest of them
            unsigned long long m = 0ull;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const unsigned long long key = i < na ? a[i] : b[i - na];
                if ((key & mask) == prefix && key > m) m = key;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long x = __shfl_xor(m, o, 64);
                m = x > m ? x : m;
            }
            if (lane == 0) red[wave] = m;
            __syncthreads();
            m = red[0];
            for (int w = 1; w < nwave; ++w) m = red[w] > m ? red[w] : m;
            __syncthreads();
            return m;
        }
    }
    return prefix;
}

// Inclusive prefix sum over the 64 lanes of a wavefront (all lanes active) with DPP moves only:
// row_shr 1, 2, 4, 8 scan each 16-lane row, then row_bcast:15 adds row 0's total to row 1 and row
// 2's to row 3, and row_bcast:31 adds the total of rows 0-1 to rows 2 and 3. Lanes whose source
// lies outside their row, and rows outside the row mask, add the old value 0. Six VALU moves
// instead of six LDS permutes.
__device__ __forceinline__ unsigned int wave_incl_scan(unsigned int x) {
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Sum over the wavefront of a per-lane count below 128, as a wave-uniform value: one ballot per
// bit of the count, no cross-lane data movement.
__device__ __forceinline__ int wave_sum_small(unsigned int x) {
    int s = 0;
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) s += __builtin_popcountll(__ballot((x >> bit) & 1u)) << bit;
    return s;
}

// A bracket (lo, hi] around the r-th smallest (1-based, 1 <= r <= 64*E, E <= 64) of the 64*E
// unsigned values a wavefront holds E per lane (callers pad with 0xffffffff, which then counts
// among the values): c_lo = count(<= lo) < r <= c_hi = count(<= hi), and either c_hi - c_lo <= 64
// or hi - lo <= 1 (every value in the bracket equals hi). A search over the value range that
// alternates between interpolation on the counts (the probe's fraction from a hardware
// reciprocal: any probe inside the bracket is valid) and bisection, so it never needs more than
// twice the bisection's steps; r = 1 is the minimum at once. r above the count of values other
// than the padding gives lo = their maximum, hi = 0xffffffff, c_lo = that count. All 64 lanes
// must call it.
struct KthBracket {
    long long lo, hi;
    int c_lo, c_hi;
};
template <int E>
__device__ __forceinline__ KthBracket wave_kth_bracket(const unsigned int (&v)[E], int r) {
    static_assert(E <= 64, "per-lane counts must stay below 128");
    // range of the values other than the padding (which would stretch it to 2^32)
    unsigned int mn = 0xffffffffu, mx = 0u, npad = 0u;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        mn = min(mn, v[e]);
        mx = v[e] == 0xffffffffu ? mx : max(mx, v[e]);
        npad += v[e] == 0xffffffffu ? 1u : 0u;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        mn = min(mn, shfl_xor_pow2(mn, o));
        mx = max(mx, shfl_xor_pow2(mx, o));
    }
    const int c_valid = 64 * E - wave_sum_small(npad);
    KthBracket b;
    b.lo = (long long)(unsigned int)__builtin_amdgcn_readfirstlane(mn) - 1;
    b.hi = (long long)(unsigned int)__builtin_amdgcn_readfirstlane(mx);
    b.c_lo = 0;
    b.c_hi = c_valid;
    if (r > c_valid) {
        b.lo = b.hi;
        b.hi = 0xffffffffll;
        b.c_lo = c_valid;
        b.c_hi = 64 * E;
        return b;
    }
    if (r == 1) b.hi = b.lo + 1;  // the minimum
    bool interp = true;
    while (b.c_hi - b.c_lo > 64 && b.hi - b.lo > 1) {
        long long mid;
        if (interp) {
            const float f = (float)(r - b.c_lo) * __builtin_amdgcn_rcpf((float)(b.c_hi - b.c_lo));
            mid = b.lo + (long long)(f * (float)(b.hi - b.lo));
            mid = mid <= b.lo ? b.lo + 1 : (mid >= b.hi ? b.hi - 1 : mid);
        } else {
            mid = b.lo + ((b.hi - b.lo) >> 1);
        }
        interp = !interp;
        const unsigned int t = (unsigned int)mid;  // lo < mid < hi, so 0 <= mid < 2^32
        unsigned int cl = 0u;
#pragma unroll
        for (int e = 0; e < E; ++e) cl += v[e] <= t ? 1u : 0u;
        const int c = wave_sum_small(cl);
        if (c >= r) {
            b.hi = mid;
            b.c_hi = c;
        } else {
            b.lo = mid;
            b.c_lo = c;
        }
    }
    return b;
}

// The r-th smallest (1-based, 1 <= r <= 64*E) of the values as above (0xffffffff when r exceeds
// the count of values other than the padding): the bracket's values, at most 64, are gathered
// (s_tmp: 64 words of LDS for this wavefront alone) and sorted across the wavefront.
template <int E>
__device__ __forceinline__ unsigned int wave_kth_u32(const unsigned int (&v)[E], int r, unsigned int* s_tmp) {
    const int lane = threadIdx.x & 63;
    const KthBracket b = wave_kth_bracket<E>(v, r);
    // every value in the bracket equals hi (or r exceeds the values: hi is the padding)
    if (b.hi - b.lo <= 1 || b.c_hi - b.c_lo > 64) return (unsigned int)b.hi;
    int base = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const bool in = (long long)v[e] > b.lo && (long long)v[e] <= b.hi;
        const unsigned long long m = __ballot(in);
        if (in) s_tmp[base + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = v[e];
        base += __builtin_popcountll(m);
    }
    __threadfence_block();  // the words other lanes stored are read below
    unsigned int x[1] = {lane < base ? s_tmp[lane] : 0xffffffffu};
    wave_bitonic_sort<1>(x);
    return (unsigned int)__shfl((int)x[0], r - b.c_lo - 1, 64);
}

// Sorts the n (<= 2048) keys produced by key_of(j), j in [0, n), with one wavefront and hands
// every element of the sorted order to emit(rank, key); ranks >= n hold HV_KEY_PAD. All 64 lanes
// of the wavefront must call it with the same n.
template <int E, typename KeyOf, typename Emit>
__device__ __forceinline__ void wave_sort_emit_e(int n, KeyOf&& key_of, Emit&& emit) {
    const int lane = threadIdx.x & 63;
    unsigned long long v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;  // any placement works: the sort orders them
        v[e] = j < n ? key_of(j) : HV_KEY_PAD;
    }
    wave_bitonic_sort<E>(v);
#pragma unroll
    for (int e = 0; e < E; ++e) emit(lane * E + e, v[e]);
}

template <typename KeyOf, typename Emit>
__device__ __forceinline__ void wave_sort_emit(int n, KeyOf&& key_of, Emit&& emit) {
    if (n <= 64) wave_sort_emit_e<1>(n, key_of, emit);
    else if (n <= 128) wave_sort_emit_e<2>(n, key_of, emit);
    else if (n <= 256) wave_sort_emit_e<4>(n, key_of, emit);
    else if (n <= 512) wave_sort_emit_e<8>(n, key_of, emit);
    else if (n <= 1024) wave_sort_emit_e<16>(n, key_of, emit);
    else wave_sort_emit_e<32>(n, key_of, emit);
}
constexpr int kWaveSortMax = 2048;

// The k smallest of n (<= kWaveSortMax) unique 64-bit (distance, id) keys handed to emit(rank,
// key) in order (ranks >= k may follow, as from wave_sort_emit), with one wavefront: for n well
// above k, the k-th smallest distance word by wave_kth_u32, then only the keys whose distance is
// <= it (every key of the top k, and at least k keys) gathered into s_buf (kTopkBuf keys of LDS
// for this wavefront) and sorted; rows with few keys, or with more ties at the k-th distance than
// the buffer holds, are sorted whole. s_tmp: 64 words of LDS for this wavefront. All 64 lanes must
// call it with the same n and k.
constexpr int kTopkBuf = 512;
template <int E, typename KeyOf, typename Emit>
__device__ __forceinline__ void wave_topk_emit_e(int n, int k, KeyOf&& key_of, unsigned int* s_tmp,
                                                 unsigned long long* s_buf, Emit&& emit) {
    const int lane = threadIdx.x & 63;
    unsigned int u[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;
        u[e] = j < n ? (unsigned int)(key_of(j) >> 32) : 0xffffffffu;
    }
    const unsigned int tau = wave_kth_u32<E>(u, k, s_tmp);
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;
        const bool in = j < n && u[e] <= tau;
        const unsigned long long m = __ballot(in);
        const int pos = cnt + (int)__builtin_popcountll(m & ((1ull << lane) - 1ull));
        if (in && pos < kTopkBuf) s_buf[pos] = key_of(j);
        cnt += (int)__builtin_popcountll(m);
    }
    if (cnt > kTopkBuf) {
        wave_sort_emit(n, key_of, emit);
        return;
    }
    __threadfence_block();  // the keys other lanes stored are read below
    wave_sort_emit(cnt, [&](int j) { return s_buf[j]; }, emit);
}

template <typename KeyOf, typename Emit>
__device__ __forceinline__ void wave_topk_emit(int n, int k, KeyOf&& key_of, unsigned int* s_tmp,
                                               unsigned long long* s_buf, Emit&& emit) {
    auto keys = [&](int j) { return j < n ? key_of(j) : HV_KEY_PAD; };
    if (n <= 256 || n <= 2 * k || 2 * k > kTopkBuf) {
        wave_sort_emit(max(n, k), keys, emit);
    } else if (n <= 512) {
        wave_topk_emit_e<8>(n, k, keys, s_tmp, s_buf, emit);
    } else if (n <= 1024) {
        wave_topk_emit_e<16>(n, k, keys, s_tmp, s_buf, emit);
    } else {
        wave_topk_emit_e<32>(n, k, keys, s_tmp, s_buf, emit);
    }
}
