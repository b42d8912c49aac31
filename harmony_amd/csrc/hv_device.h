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
