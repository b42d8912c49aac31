// Distance tiles, row selects and index-build kernels for the IVF-flat path (gfx950).
//
// Distances are direct-difference L2 in FP32 accumulated as a fused multiply-add chain in
// dimension order, acc = fma(x_i - y_i, x_i - y_i, acc): the loop of
// calculatedEuclideanDistance (src/utils.h:480-487) as a compiler that contracts it emits.
// The chain is kept per (row, column) pair, so every distance is bit-identical to the same
// chain on the host.
#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

// ---------------------------------------------------------------------------------------
// Full-d L2 between every row of Q [nq][d] and every row of C [nc][d] -> out[nq][ldo].
// 64x64 output tile per workgroup, 4x4 per thread, dims staged 32 at a time through LDS
// (transposed, so each thread reads its 4 rows / 4 columns with one 16-byte LDS load).
// Used for the coarse quantizer (Index.cpp:967-997), k-means assignment
// (Clustering.cpp:102-120) and add-time assignment (Index.cpp:596-614).
__global__ __launch_bounds__(256) void k_pair_l2_full(const float* __restrict__ Q, int64_t nq,
                                                      const float* __restrict__ C, int64_t nc, int d,
                                                      float* __restrict__ out, int64_t ldo) {
    constexpr int TQ = 64, TC = 64, KC = 32, LS = 68;
    __shared__ __align__(16) float Qs[KC][LS];
    __shared__ __align__(16) float Cs[KC][LS];
    const int tid = threadIdx.x;
    const int64_t q0 = (int64_t)blockIdx.y * TQ, c0 = (int64_t)blockIdx.x * TC;
    const int tq = tid >> 4, tc = tid & 15;
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
                if (q0 + r < nq) v = Q[(q0 + r) * d + k0 + kk];
                if (c0 + r < nc) w = C[(c0 + r) * d + k0 + kk];
            }
            Qs[kk][r] = v;
            Cs[kk][r] = w;
        }
        __syncthreads();
        for (int kk = 0; kk < kn; ++kk) {
            const float4 a = *reinterpret_cast<const float4*>(&Qs[kk][tq * 4]);
            const float4 b = *reinterpret_cast<const float4*>(&Cs[kk][tc * 4]);
            const float qa[4] = {a.x, a.y, a.z, a.w};
            const float cb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float t = qa[i] - cb[j];
                    acc[i][j] = __builtin_fmaf(t, t, acc[i][j]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t r = q0 + tq * 4 + i;
        if (r >= nq) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t c = c0 + tc * 4 + j;
            if (c < nc) out[r * ldo + c] = acc[i][j];
        }
    }
}

void launch_pair_l2_full(const float* Q, int64_t nq, const float* C, int64_t nc, int d, float* out, int64_t ldo,
                         hipStream_t s) {
    if (nq <= 0 || nc <= 0) return;
    dim3 grid((unsigned)((nc + 63) / 64), (unsigned)((nq + 63) / 64));
    hipLaunchKernelGGL(k_pair_l2_full, grid, dim3(256), 0, s, Q, nq, C, nc, d, out, ldo);
}

// ---------------------------------------------------------------------------------------
// Per row: argmin over (distance, column) — nearest centroid with ties to the lower id.
// One wavefront per row, 64-bit key min-reduction across the wavefront.
__global__ __launch_bounds__(256) void k_argmin_rows(const float* __restrict__ dist, int64_t n, int m,
                                                     int* __restrict__ labels, float* __restrict__ best) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    unsigned long long kmin = HV_KEY_PAD;
    for (int c = lane; c < m; c += 64) {
        const unsigned long long k = hv_key(dist[row * m + c], (uint32_t)c);
        kmin = k < kmin ? k : kmin;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(kmin, off, 64);
        kmin = o < kmin ? o : kmin;
    }
    if (lane == 0) {
        labels[row] = (int)(kmin & 0xffffffffu);
        if (best) best[row] = hv_key_dist(kmin);
    }
}

void launch_argmin_rows(const float* dist, int64_t n, int m, int* labels, float* best, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_argmin_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, dist, n, m, labels, best);
}

// ---------------------------------------------------------------------------------------
// dst[r][i] = src[r][perm[i]]: applies the index's dimension order to incoming vectors. One
// workgroup per row (no division per element), the row's gathers served by the cache.
__global__ void k_permute_rows(const float* __restrict__ src, int64_t n, int d, const int* __restrict__ perm,
                               float* __restrict__ dst) {
    const int64_t r = blockIdx.x;
    const float* __restrict__ sr = src + r * d;
    float* __restrict__ dr = dst + r * d;
    for (int i = threadIdx.x; i < d; i += blockDim.x) dr[i] = sr[perm[i]];
}

void launch_permute_rows(const float* src, int64_t n, int d, const int* perm, float* dst, hipStream_t s) {
    if (n <= 0 || d <= 0) return;
    hipLaunchKernelGGL(k_permute_rows, dim3((unsigned)n), dim3(d <= 128 ? 128 : 256), 0, s, src, n, d, perm, dst);
}

// ---------------------------------------------------------------------------------------
// Fill kernels (used instead of hipMemsetAsync so a captured graph never holds a memset).
__global__ void k_fill_i32(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
// The segments share one grid-stride loop each (a search clears ~7 small buffers at once).
__global__ void k_fill_multi(FillSegs f) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = 0; j < f.nseg; ++j) {
        int* p = f.p[j];
        const int64_t n = f.n[j];
        const int v = f.v[j];
        for (int64_t i = g; i < n; i += stride) p[i] = v;
    }
}
void launch_fill_multi(const FillSegs& f, hipStream_t s) {
    int64_t mx = 0;
    for (int j = 0; j < f.nseg; ++j) mx = std::max<int64_t>(mx, f.n[j]);
    if (mx <= 0) return;
    const int64_t blocks = std::min<int64_t>((mx + 255) / 256, 2048);
    hipLaunchKernelGGL(k_fill_multi, dim3((unsigned)blocks), dim3(256), 0, s, f);
}

// ---------------------------------------------------------------------------------------
// k-means centroid update with FP64 sums (Clustering.cpp:93-152, sums in FP64 instead of the
// reference's FP32 omp-critical order): perm lists the sample points grouped by cluster in
// ascending point order, cl_off[nlist+1] delimits the clusters. Each thread owns dimensions
// and sums its cluster's points in that fixed order, so the result is deterministic and equal
// to a sequential host loop. Empty clusters keep their centroid (Clustering.cpp:146-150).
__global__ __launch_bounds__(256) void k_centroid_update_f64(const float* __restrict__ x, int d,
                                                             const int* __restrict__ perm,
                                                             const int64_t* __restrict__ cl_off,
                                                             float* __restrict__ centroids) {
    const int c = blockIdx.x;
    const int64_t b = cl_off[c], e = cl_off[c + 1];
    if (e == b) return;
    const double cnt = (double)(e - b);
    for (int t = threadIdx.x; t < d; t += blockDim.x) {
        double s = 0.0;
        for (int64_t p = b; p < e; ++p) s += (double)x[(int64_t)perm[p] * d + t];
        centroids[(int64_t)c * d + t] = (float)(s / cnt);
    }
}

void launch_centroid_update_f64(const float* x, int d, const int* perm, const int64_t* cl_off, int nlist,
                                float* centroids, hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_centroid_update_f64, dim3((unsigned)nlist), dim3(256), 0, s, x, d, perm, cl_off,
                       centroids);
}

// ---------------------------------------------------------------------------------------
// List-major, dimension-major layout: list l occupies codes[code_off[l] .. + d*stride_l),
// row i (dimension i in index order) holds the list's vectors in within-list order, padded
// with zeros to stride_l (a multiple of 16 floats). A dimension block of a list is therefore
// one contiguous slab, and a tile of consecutive vectors of one dimension is one contiguous,
// 64-byte aligned run.
__global__ __launch_bounds__(256) void k_build_layout(const float* __restrict__ x, int d,
                                                      const int* __restrict__ order,
                                                      const int64_t* __restrict__ list_off,
                                                      const int64_t* __restrict__ code_off,
                                                      const int* __restrict__ list_stride,
                                                      float* __restrict__ codes) {
    const int l = blockIdx.x;
    const int64_t b = list_off[l];
    const int n = (int)(list_off[l + 1] - b);
    const int ps = list_stride[l];
    float* dst = codes + code_off[l];
    for (int i = 0; i < d; ++i)
        for (int t = threadIdx.x; t < ps; t += blockDim.x)
            dst[(int64_t)i * ps + t] = t < n ? x[(int64_t)order[b + t] * d + i] : 0.f;
}

void launch_build_layout(const float* x, int d, const int* order, const int64_t* list_off,
                         const int64_t* code_off, const int* list_stride, int nlist, float* codes, hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_build_layout, dim3((unsigned)nlist), dim3(256), 0, s, x, d, order, list_off, code_off,
                       list_stride, codes);
}

}  // namespace hv
