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

__host__ __device__ inline int hv_next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Wavefront-wide ascending bitonic sort of 64*E keys held in registers: element i lives in
// lane i / E, slot i % E. Strides below E exchange within a lane; larger strides exchange with
// lane ^ (stride / E) through cross-lane shuffles. No LDS, no barriers.
template <int E>
__device__ __forceinline__ void wave_bitonic_sort(unsigned long long (&v)[E]) {
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
                    const unsigned long long o = __shfl_xor(v[e], lx, 64);
                    const bool up = (i & size) == 0;
                    const bool lower = (i & stride) == 0;
                    const unsigned long long mn = o < v[e] ? o : v[e];
                    const unsigned long long mx = o < v[e] ? v[e] : o;
                    v[e] = (lower == up) ? mn : mx;
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & stride) continue;
                    const int e2 = e | stride;
                    const bool up = ((lane * E + e) & size) == 0;
                    const unsigned long long a = v[e], b = v[e2];
                    const bool sw = (a > b) == up;
                    v[e] = sw ? b : a;
                    v[e2] = sw ? a : b;
                }
            }
        }
    }
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
