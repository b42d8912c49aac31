// Inner-product metric (METRIC_IP): the reference's IP training (Clustering.cpp:121-151 with the
// centroid normalization of Index.cpp:408-418), assignment and list order (Index.cpp:567-671:
// each vector to the centroid with the largest inner product, lists by descending inner product
// with their centroid, the candidate2centroid array) and the coarse step of its ORIGINAL search
// (Index.cpp:1440-1448 with the IP scanner). The list scan with triangle pruning is the exact
// path's IP instantiation (hv_scan.hip). Every inner product is the FP32 FMA chain in dimension
// order, as utils.h:502-520 computes it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

// out[r][c] = <X_r, C_c> for a 64 x 64 tile per workgroup, 32-dim chunks staged in LDS, each
// thread 4 x 4 chains.
__global__ __launch_bounds__(256) void k_ip_rows(const float* __restrict__ X, int64_t n, const float* __restrict__ C,
                                                 int64_t nc, int d, float* __restrict__ out, int64_t ldo) {
    constexpr int TQ = 64, TC = 64, KC = 32, LS = 68;
    __shared__ __align__(16) float Xs[KC][LS];
    __shared__ __align__(16) float Cs[KC][LS];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * TQ, c0 = (int64_t)blockIdx.x * TC;
    const int tr = tid >> 4, tc = tid & 15;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k0 = 0; k0 < d; k0 += KC) {
        const int kn = min(KC, d - k0);
        for (int e = tid; e < TQ * KC; e += 256) {
            const int r = e & 63, kk = e >> 6;
            float v = 0.f, w = 0.f;
            if (kk < kn) {
                if (r0 + r < n) v = X[(r0 + r) * d + k0 + kk];
                if (c0 + r < nc) w = C[(c0 + r) * d + k0 + kk];
            }
            Xs[kk][r] = v;
            Cs[kk][r] = w;
        }
        __syncthreads();
        for (int kk = 0; kk < kn; ++kk) {
            const float4 a = *reinterpret_cast<const float4*>(&Xs[kk][tr * 4]);
            const float4 b = *reinterpret_cast<const float4*>(&Cs[kk][tc * 4]);
            const float xa[4] = {a.x, a.y, a.z, a.w};
            const float cb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf(xa[i], cb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t r = r0 + tr * 4 + i;
        if (r >= n) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t c = c0 + tc * 4 + j;
            if (c < nc) out[r * ldo + c] = acc[i][j];
        }
    }
}

void launch_ip_rows(const float* X, int64_t n, const float* C, int64_t nc, int d, float* out, int64_t ldo,
                    hipStream_t s) {
    if (n <= 0 || nc <= 0) return;
    // row tiles run on grid.y, at most 65535 of them per launch
    constexpr int64_t kRowsPerLaunch = 65535 * 64;
    for (int64_t r0 = 0; r0 < n; r0 += kRowsPerLaunch) {
        const int64_t rows = n - r0 < kRowsPerLaunch ? n - r0 : kRowsPerLaunch;
        const dim3 grid((unsigned)((nc + 63) / 64), (unsigned)((rows + 63) / 64));
        hipLaunchKernelGGL(k_ip_rows, grid, dim3(256), 0, s, X + r0 * d, rows, C, nc, d, out + r0 * ldo, ldo);
    }
}

// Per row: the column with the largest value, ties to the lower column (a strict '>' scan in
// column order), and that value. One wavefront per row.
__global__ __launch_bounds__(256) void k_argmax_rows(const float* __restrict__ v, int64_t n, int m,
                                                     int* __restrict__ labels, float* __restrict__ best) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    unsigned long long kmin = HV_IP_KEY_PAD;
    for (int c = lane; c < m; c += 64) {
        const unsigned long long k = hv_ip_key(v[row * m + c], (uint32_t)c);
        kmin = k < kmin ? k : kmin;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(kmin, off, 64);
        kmin = o < kmin ? o : kmin;
    }
    if (lane == 0) {
        labels[row] = (int)(kmin & 0xffffffffu);
        if (best) best[row] = v[row * m + (kmin & 0xffffffffu)];
    }
}

void launch_argmax_rows(const float* v, int64_t n, int m, int* labels, float* best, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_argmax_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, v, n, m, labels, best);
}

// out_r = x_r / sqrt(<x_r, x_r>) (a zero row is copied), the norm from the FP32 chain; one
// wavefront per row, lane 0 forms the chain in dimension order. sqrtf and '/' are the correctly
// rounded IEEE operations under hipcc's default -fhip-fp32-correctly-rounded-divide-sqrt (HIP's
// __fsqrt_rn is the approximate native square root unless OCML_BASIC_ROUNDED_OPERATIONS is set).
__global__ __launch_bounds__(256) void k_normalize_rows(const float* __restrict__ x, int64_t n, int d,
                                                        float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    float nn = 0.f;
    if (lane == 0) {
        float acc = 0.f;
        for (int i = 0; i < d; ++i) acc = __builtin_fmaf(x[row * d + i], x[row * d + i], acc);
        nn = sqrtf(acc);
    }
    nn = __shfl(nn, 0, 64);
    for (int i = lane; i < d; i += 64) out[row * d + i] = nn > 0.f ? x[row * d + i] / nn : x[row * d + i];
}

void launch_normalize_rows(const float* x, int64_t n, int d, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_normalize_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, x, n, d, out);
}

// In place: c_l /= sqrt(<c_l, c_l>) when that norm is not 0 (Index.cpp:408-418).
__global__ __launch_bounds__(256) void k_normalize_centroids(float* __restrict__ c, int nlist, int d) {
    const int lane = threadIdx.x & 63;
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (l >= nlist) return;
    float r = 0.f;
    if (lane == 0) {
        float acc = 0.f;
        for (int i = 0; i < d; ++i) acc = __builtin_fmaf(c[(int64_t)l * d + i], c[(int64_t)l * d + i], acc);
        r = acc != 0.f ? sqrtf(acc) : 0.f;
    }
    r = __shfl(r, 0, 64);
    if (r != 0.f)
        for (int i = lane; i < d; i += 64) c[(int64_t)l * d + i] = c[(int64_t)l * d + i] / r;
}

void launch_normalize_centroids(float* c, int nlist, int d, hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_normalize_centroids, dim3((unsigned)((nlist + 3) / 4)), dim3(256), 0, s, c, nlist, d);
}

// The P largest of each row (inner product descending, column ascending), with their values:
// one workgroup per row sorts the row's keys in LDS (m <= kMaxSelectRow).
__global__ __launch_bounds__(256) void k_ip_select_rows(const float* __restrict__ v, int64_t n, int m, int P,
                                                        int* __restrict__ ids, float* __restrict__ vals) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t row = blockIdx.x;
    const int M = hv_next_pow2(max(m, 2));
    for (int i = threadIdx.x; i < M; i += blockDim.x)
        s_keys[i] = i < m ? hv_ip_key(v[row * m + i], (uint32_t)i) : HV_IP_KEY_PAD;
    block_bitonic_sort(s_keys, M);
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const unsigned long long key = s_keys[i];
        ids[row * P + i] = (int)(key & 0xffffffffu);
        vals[row * P + i] = hv_ip_key_value(key);
    }
}

void launch_ip_select_rows(const float* v, int64_t n, int m, int P, int* ids, float* vals, hipStream_t s) {
    if (n <= 0) return;
    const size_t lds = sizeof(unsigned long long) * hv_next_pow2(m > 2 ? m : 2);
    hipLaunchKernelGGL(k_ip_select_rows, dim3((unsigned)n), dim3(256), lds, s, v, n, m, P, ids, vals);
}

// simi[qlist[q0 + i]] = the k-th inner product of that query's current result row, or -FLT_MAX
// when the row holds fewer than k results (the IP heap's lowest, heap.hpp:13-25).
__global__ void k_ip_simi(const int* __restrict__ qlist, const int* __restrict__ nql_dev, int q0, int nql_max,
                          const float* __restrict__ D, const int64_t* __restrict__ I, int k, float* __restrict__ simi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nql = min(*nql_dev - q0, nql_max);
    if (i >= nql) return;
    const int64_t q = qlist[q0 + i];
    simi[q] = I[q * k + k - 1] >= 0 ? D[q * k + k - 1] : -3.402823466e38f;
}

void launch_ip_simi(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* D, const int64_t* I, int k,
                    float* simi, hipStream_t s) {
    if (nql_max <= 0) return;
    hipLaunchKernelGGL(k_ip_simi, dim3((unsigned)((nql_max + 255) / 256)), dim3(256), 0, s, qlist, nql_dev, q0,
                       nql_max, D, I, k, simi);
}

// list[0] = n, list[1 + i] = i
__global__ void k_iota_list(int* __restrict__ list, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) list[0] = (int)n;
    if (i < n) list[1 + i] = (int)i;
}

void launch_iota_list(int* list, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_iota_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, list, n);
}

}  // namespace hv
