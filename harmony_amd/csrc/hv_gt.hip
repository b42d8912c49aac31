// FP64 brute-force ground truth over every vector of the index (the exact kNN that replaces
// the reference's faiss IndexIVFFlat at nprobe = nlist, query.cpp:481-514). Each query comes
// with an upper bound tau on its k-th FP64 distance; the kernel computes the FP64 distance
// (q_i - x_i)^2 summed in double for every (query, vector) pair and keeps the pairs with
// distance <= tau. The host sorts the few survivors by (distance, id).
#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

// grid.x = list, grid.y = tile of 64 queries; each thread owns 4 queries x 4 vectors.
__global__ __launch_bounds__(256) void k_gt_f64(const float* __restrict__ Qp, int64_t nq, int d,
                                                const float* __restrict__ codes, const int64_t* __restrict__ code_off,
                                                const int64_t* __restrict__ list_off,
                                                const int* __restrict__ list_stride, const uint32_t* __restrict__ ids,
                                                const double* __restrict__ tau, int cap, double* __restrict__ sd,
                                                uint32_t* __restrict__ sid, unsigned int* __restrict__ cnt) {
    constexpr int TQ = 64, TV = 64, KC = 32, LS = 68;
    __shared__ __align__(16) float Qs[KC][LS];
    __shared__ __align__(16) float Xs[KC][LS];
    const int l = blockIdx.x;
    const int64_t q0 = (int64_t)blockIdx.y * TQ;
    const int64_t lb = list_off[l];
    const int n = (int)(list_off[l + 1] - lb);
    const int ps = list_stride[l];
    const float* __restrict__ X = codes + code_off[l];
    const int tid = threadIdx.x, tq = tid >> 4, tv = tid & 15;
    double myt[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) myt[i] = (q0 + tq * 4 + i < nq) ? tau[q0 + tq * 4 + i] : -1.0;
    for (int v0 = 0; v0 < n; v0 += TV) {
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int k0 = 0; k0 < d; k0 += KC) {
            const int kn = min(KC, d - k0);
            for (int e = tid; e < TQ * KC; e += 256) {
                const int r = e & 63, kk = e >> 6;
                float a = 0.f, b = 0.f;
                if (kk < kn) {
                    if (q0 + r < nq) a = Qp[(q0 + r) * d + k0 + kk];
                    if (v0 + r < ps) b = X[(int64_t)(k0 + kk) * ps + v0 + r];
                }
                Qs[kk][r] = a;
                Xs[kk][r] = b;
            }
            __syncthreads();
            for (int kk = 0; kk < kn; ++kk) {
                const float4 a = *reinterpret_cast<const float4*>(&Qs[kk][tq * 4]);
                const float4 b = *reinterpret_cast<const float4*>(&Xs[kk][tv * 4]);
                const double qa[4] = {a.x, a.y, a.z, a.w};
                const double xb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double t = qa[i] - xb[j];
                        acc[i][j] = __builtin_fma(t, t, acc[i][j]);
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int v = v0 + tv * 4 + j;
                if (v < n && acc[i][j] <= myt[i]) {
                    const int64_t q = q0 + tq * 4 + i;
                    const unsigned int slot = atomicAdd(&cnt[q], 1u);
                    if (slot < (unsigned int)cap) {
                        sd[q * cap + slot] = acc[i][j];
                        sid[q * cap + slot] = ids[lb + v];
                    }
                }
            }
    }
}

void launch_gt_f64(const float* Qp, int64_t nq, int d, const float* codes, const int64_t* code_off,
                   const int64_t* list_off, const int* list_stride, const uint32_t* ids, int nlist,
                   const double* tau, int cap, double* sd, uint32_t* sid, unsigned int* cnt, hipStream_t s) {
    if (nq <= 0) return;
    dim3 grid((unsigned)nlist, (unsigned)((nq + 63) / 64));
    hipLaunchKernelGGL(k_gt_f64, grid, dim3(256), 0, s, Qp, nq, d, codes, code_off, list_off, list_stride, ids, tau,
                       cap, sd, sid, cnt);
}

}  // namespace hv
