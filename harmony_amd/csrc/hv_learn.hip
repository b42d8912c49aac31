// Learned probe/bound predictor: a 65 -> 128 (ReLU) -> 2 MLP over a query's distances to its 64
// nearest centroids and its squared norm, predicting log1p(k-th neighbour distance) and
// log(nprobe needed for the target recall).
//
// Features (raw): f_i = log1p(i-th smallest coarse distance bound), i < 64 (the last one repeated
// when there are fewer lists), f_64 = log1p(|q|^2). The bounds are the coarse step's upper bounds
// d~ + kappa (|q|^2 + |c|^2) of the FP32 chain distances (d~ the MFMA estimate), sorted: the
// search's probe selection forms them on the way (hv_coarse.hip). The network works on standardized features
// xn = (f - fmean) / fstd and standardized targets; outputs are mapped back with the target
// mean/std. Training keeps FP32 master weights (forward, backward and Adam below, all FP32);
// serving uses BF16 copies of the weight matrices with FP32 accumulation.
//
// Parameter layout (flat FP32): W1 [65][128] (input-major), b1 [128], W2 [128][2], b2 [2].
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hv_kernels.h"

namespace hv {

constexpr int kIn = kMlpIn, kHid = kMlpHidden, kOut = kMlpOut;
constexpr int kOffB1 = kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kOut;
constexpr int kInPad = (kIn + 3) / 4 * 4;  // the serving copy's inputs, in groups of 4
static_assert(kOffB2 + kOut == kMlpParams, "parameter layout");

// ---------------------------------------------------------------------------------------
// Features from the sorted coarse distance estimates (pdist [n][P], ascending) and the queries.
__global__ void k_features(const float* __restrict__ pdist, int P, const float* __restrict__ Qp, int d, int64_t n,
                           float* __restrict__ feat) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= n) return;
    const int lane = threadIdx.x & 63;
    float s = 0.f;
    for (int i = lane; i < d; i += 64) {
        const float v = Qp[q * d + i];
        s = __builtin_fmaf(v, v, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const int c = min(lane, P - 1);
    feat[q * kIn + lane] = log1pf(fmaxf(pdist[q * P + c], 0.f));
    if (lane == 0) feat[q * kIn + kNearest] = log1pf(s);
}

void launch_features(const float* pdist, int P, const float* Qp, int d, int64_t n, float* feat, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, pdist, P, Qp, d, n, feat);
}

// ---------------------------------------------------------------------------------------
// Training. A batch is rows idx[0..B) of X [N][65] / T [N][2] (raw features / raw targets).
// Forward: one workgroup of 128 threads per row; thread j owns hidden unit j.
__global__ __launch_bounds__(128) void k_mlp_fwd(const float* __restrict__ X, const int* __restrict__ idx,
                                                 const float* __restrict__ P, const float* __restrict__ norm,
                                                 float* __restrict__ Xn, float* __restrict__ H,
                                                 float* __restrict__ Y) {
    __shared__ float xs[kIn];
    __shared__ float hs[kHid];
    const int b = blockIdx.x, j = threadIdx.x;
    const int64_t r = idx[b];
    const float* fmean = norm;
    const float* fstd = norm + kIn;
    if (j < kIn) {
        const float v = (X[r * kIn + j] - fmean[j]) / fstd[j];
        xs[j] = v;
        Xn[(int64_t)b * kIn + j] = v;
    }
    __syncthreads();
    float h = P[kOffB1 + j];
    for (int i = 0; i < kIn; ++i) h = __builtin_fmaf(xs[i], P[i * kHid + j], h);
    h = fmaxf(h, 0.f);
    hs[j] = h;
    H[(int64_t)b * kHid + j] = h;
    __syncthreads();
    if (j < kOut) {
        float y = P[kOffB2 + j];
        for (int u = 0; u < kHid; ++u) y = __builtin_fmaf(hs[u], P[kOffW2 + u * kOut + j], y);
        Y[(int64_t)b * kOut + j] = y;
    }
}

// Output and hidden gradients of L = 1/(2B) sum_b sum_o (y_bo - t_bo)^2 (standardized targets),
// and each row's loss term.
__global__ __launch_bounds__(128) void k_mlp_bwd_rows(const float* __restrict__ T, const int* __restrict__ idx,
                                                      int B, const float* __restrict__ P,
                                                      const float* __restrict__ norm, const float* __restrict__ H,
                                                      const float* __restrict__ Y, float* __restrict__ dY,
                                                      float* __restrict__ dH, float* __restrict__ row_loss) {
    __shared__ float dys[kOut];
    const int b = blockIdx.x, j = threadIdx.x;
    const float* tmean = norm + 2 * kIn;
    const float* tstd = tmean + kOut;
    if (j < kOut) {
        const float t = (T[(int64_t)idx[b] * kOut + j] - tmean[j]) / tstd[j];
        const float e = Y[(int64_t)b * kOut + j] - t;
        dys[j] = e / (float)B;
        dY[(int64_t)b * kOut + j] = e / (float)B;
    }
    __syncthreads();
    if (j == 0) {
        const float e0 = dys[0] * (float)B, e1 = dys[1] * (float)B;
        row_loss[b] = 0.5f * (e0 * e0 + e1 * e1) / (float)B;
    }
    float g = 0.f;
    for (int o = 0; o < kOut; ++o) g = __builtin_fmaf(dys[o], P[kOffW2 + j * kOut + o], g);
    dH[(int64_t)b * kHid + j] = H[(int64_t)b * kHid + j] > 0.f ? g : 0.f;
}

// Parameter gradients: one thread per parameter, summing its batch terms in row order.
__global__ void k_mlp_bwd_params(int B, const float* __restrict__ Xn, const float* __restrict__ H,
                                 const float* __restrict__ dY, const float* __restrict__ dH, float* __restrict__ G) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = 0.f;
    if (p < kOffB1) {
        const int i = p / kHid, j = p % kHid;
        for (int b = 0; b < B; ++b) g = __builtin_fmaf(Xn[(int64_t)b * kIn + i], dH[(int64_t)b * kHid + j], g);
    } else if (p < kOffW2) {
        const int j = p - kOffB1;
        for (int b = 0; b < B; ++b) g += dH[(int64_t)b * kHid + j];
    } else if (p < kOffB2) {
        const int j = (p - kOffW2) / kOut, o = (p - kOffW2) % kOut;
        for (int b = 0; b < B; ++b) g = __builtin_fmaf(H[(int64_t)b * kHid + j], dY[(int64_t)b * kOut + o], g);
    } else {
        const int o = p - kOffB2;
        for (int b = 0; b < B; ++b) g += dY[(int64_t)b * kOut + o];
    }
    G[p] = g;
}

// Sum of the rows' losses (one workgroup, fixed order) added to *acc.
__global__ __launch_bounds__(256) void k_sum_loss(const float* __restrict__ row_loss, int B, float* __restrict__ acc) {
    __shared__ float s[256];
    float v = 0.f;
    for (int b = threadIdx.x; b < B; b += 256) v += row_loss[b];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += s[0];
}

// Adam (Kingma & Ba 2015) with bias correction; t is the 1-based step.
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M, float* __restrict__ V,
                       float lr, float beta1, float beta2, float eps, int t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = G[p];
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    const float mh = m / (1.f - powf(beta1, (float)t));
    const float vh = v / (1.f - powf(beta2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

void launch_mlp_step(const float* X, const float* T, const int* idx, int B, float* P, const float* norm, float* Xn,
                     float* H, float* Y, float* dY, float* dH, float* row_loss, float* G, float* loss_acc,
                     hipStream_t s) {
    if (B <= 0) return;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)B), dim3(kHid), 0, s, X, idx, P, norm, Xn, H, Y);
    hipLaunchKernelGGL(k_mlp_bwd_rows, dim3((unsigned)B), dim3(kHid), 0, s, T, idx, B, P, norm, H, Y, dY, dH,
                       row_loss);
    hipLaunchKernelGGL(k_mlp_bwd_params, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, B, Xn, H, dY, dH, G);
    if (loss_acc) hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, s, row_loss, B, loss_acc);
}

void launch_adam(float* P, const float* G, float* M, float* V, float lr, float beta1, float beta2, float eps, int t,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, P, G, M, V, lr, beta1, beta2, eps, t);
}

// ---------------------------------------------------------------------------------------
// Serving weights: BF16 copies of W1 and W2 (biases stay FP32).
// The serving copies: W1 in groups of 4 inputs (kInPad, the last padded with zeros), lane j of a
// group holding hidden units 2j and 2j + 1: element [g][j][k][h] = W1[4g + k][2j + h], so one
// 16-byte load gives a lane its 8 weights of a group and a wavefront's load of a group is one
// contiguous 1 KB; W2 as it is, [hidden][output].
__global__ void k_mlp_to_bf16(const float* __restrict__ P, __bf16* __restrict__ w1, __bf16* __restrict__ w2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpW1Serve) {
        const int g = p >> 9, j = (p >> 3) & 63, k = (p >> 1) & 3, hh = p & 1;
        const int i = 4 * g + k, u = 2 * j + hh;
        w1[p] = (__bf16)(i < kIn ? P[i * kHid + u] : 0.f);
    }
    if (p < kHid * kOut) w2[p] = (__bf16)P[kOffW2 + p];
}

void launch_mlp_to_bf16(const float* P, __bf16* w1, __bf16* w2, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((kMlpW1Serve + 255) / 256), dim3(256), 0, s, P, w1, w2);
}

// One wavefront per row: lane j owns hidden units 2j and 2j + 1 (the serving layout of W1, one
// 16-byte load per group of 4 inputs); the two outputs are reduced across the wavefront. Returns
// the outputs in raw target space (every lane gets them). The standardized inputs are formed
// once, one per lane (lanes 0..3 also the last and the padding zeros), into xs (kInPad floats of
// LDS for this wavefront), instead of by every lane for every input.
__device__ __forceinline__ void mlp_eval_bf16(const float* __restrict__ f, int lane, const __bf16* __restrict__ w1,
                                              const __bf16* __restrict__ w2, const float* __restrict__ P,
                                              const float* __restrict__ norm, float* xs, float& y0, float& y1) {
    const float* fmean = norm;
    const float* fstd = norm + kIn;
    const float* tmean = norm + 2 * kIn;
    const float* tstd = tmean + kOut;
    static_assert(kInPad <= 128 && kHid == 128 && kOut == 2, "two inputs and two hidden units per lane");
    xs[lane] = (f[lane] - fmean[lane]) / fstd[lane];
    if (lane + 64 < kIn)
        xs[lane + 64] = (f[lane + 64] - fmean[lane + 64]) / fstd[lane + 64];
    else if (lane + 64 < kInPad)
        xs[lane + 64] = 0.f;  // padding inputs (their weights are 0 too)
    __threadfence_block();  // every lane reads every input below
    const float2 b1 = *reinterpret_cast<const float2*>(P + kOffB1 + 2 * lane);
    float h0 = b1.x, h1 = b1.y;
    // each hidden unit sums its inputs in order; the 17 group loads are independent
    const uint4* __restrict__ wq = reinterpret_cast<const uint4*>(w1) + lane;
#pragma unroll
    for (int g = 0; g < kInPad / 4; ++g) {
        const uint4 wv = wq[g * 64];
        const unsigned int wk[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float x = xs[4 * g + k];
            h0 = __builtin_fmaf(x, __uint_as_float(wk[k] << 16), h0);          // unit 2j
            h1 = __builtin_fmaf(x, __uint_as_float(wk[k] & 0xffff0000u), h1);  // unit 2j + 1
        }
    }
    h0 = fmaxf(h0, 0.f);
    h1 = fmaxf(h1, 0.f);
    // W2 rows 2j and 2j + 1 (outputs 0, 1 each): 4 BF16 values, one 8-byte load
    const uint2 w2v = *reinterpret_cast<const uint2*>(w2 + 4 * lane);
    float a = h0 * __uint_as_float(w2v.x << 16) + h1 * __uint_as_float(w2v.y << 16);
    float c = h0 * __uint_as_float(w2v.x & 0xffff0000u) + h1 * __uint_as_float(w2v.y & 0xffff0000u);
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    y0 = (a + P[kOffB2]) * tstd[0] + tmean[0];
    y1 = (c + P[kOffB2 + 1]) * tstd[1] + tmean[1];
}

__global__ __launch_bounds__(256) void k_mlp_predict(const float* __restrict__ X, int64_t n,
                                                     const __bf16* __restrict__ w1, const __bf16* __restrict__ w2,
                                                     const float* __restrict__ P, const float* __restrict__ norm,
                                                     float* __restrict__ out) {
    __shared__ float s_x[4][kInPad];
    const int w = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= n) return;
    const int lane = threadIdx.x & 63;
    float y0, y1;
    mlp_eval_bf16(X + q * kIn, lane, w1, w2, P, norm, s_x[w], y0, y1);
    if (lane == 0) {
        out[q * 2] = y0;
        out[q * 2 + 1] = y1;
    }
}

void launch_mlp_predict(const float* X, int64_t n, const __bf16* w1, const __bf16* w2, const float* P,
                        const float* norm, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, X, n, w1, w2, P, norm, out);
}

// Serving: per query, the probe count qn = clamp(ceil(exp(y1) * nprobe_scale), min_nprobe,
// nprobe), raised (with list_off) to the shortest prefix of the probe list holding kmin vectors,
// and the seeded bound R = expm1(y0) * bound_scale (FLT_MAX when bound_scale <= 0), also into
// bound2 and bound3 when given (the search's working and reference bounds start there).
// The features are formed here, in LDS, from the query's kNearest smallest coarse distance
// estimates (near, ascending) and the query, with k_features' arithmetic (so the same bits);
// pdist / probes (pc per query) are the exact probe lists the margin test walks.
//
// Margin widening (margin > 0, with a finite R): a list l past the predicted count can hold a
// vector within sqrt(R) of q only if both lower bounds of |q - x| over its vectors x are <= sqrt(R):
//  - the triangle bound |q - c_l| - rad_l (rad_l = the largest |x - c_l| of the list);
//  - the Voronoi bound (|q - c_l|^2 - |q - c_0|^2) / (2 |c_l - c_0|), c_0 = the query's nearest
//    centroid: every x of list l is at least as close to c_l as to c_0 (vectors are filed under
//    their nearest centroid), so it lies beyond the bisector of c_0 and c_l, whose distance from
//    q is that quotient.
// qn grows to cover every list of the first nprobe whose larger bound g satisfies g <= 0 or
// g^2 <= R * margin, with relative slack for the FP32 roundings of the distances, norms and
// square roots involved. With margin >= 1, and once the search has checked that the query's k-th
// result is <= R, no list left out holds a vector of the top k over all nprobe lists: the query's
// result is the plain search's at nprobe. Queries whose count the margin test grew are counted in
// *widened.
__global__ __launch_bounds__(256) void k_mlp_serve(const float* __restrict__ near, const float* __restrict__ pdist,
                                                   const int* __restrict__ probes,
                                                   int pc, const float* __restrict__ Qp, int d, int64_t n,
                                                   const __bf16* __restrict__ w1, const __bf16* __restrict__ w2,
                                                   const float* __restrict__ P, const float* __restrict__ norm,
                                                   float nprobe_scale, float bound_scale, int min_nprobe, int nprobe,
                                                   const float* __restrict__ lrad, const float* __restrict__ cent,
                                                   float margin, float slack, const int64_t* __restrict__ list_off,
                                                   int kmin,
                                                   int* __restrict__ qn, float* __restrict__ bound,
                                                   float* __restrict__ pred, float* __restrict__ bound2,
                                                   float* __restrict__ bound3, unsigned long long* __restrict__ widened,
                                                   int* __restrict__ bucket_cnt, int nranks) {
    __shared__ float s_f[4][kIn];
    __shared__ float s_x[4][kInPad];
    const int w = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= n) return;
    const int lane = threadIdx.x & 63;
    float s = 0.f;
    for (int i = lane; i < d; i += 64) {
        const float v = Qp[q * d + i];
        s = __builtin_fmaf(v, v, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    s_f[w][lane] = log1pf(fmaxf(near[q * kNearest + lane], 0.f));
    if (lane == 0) s_f[w][kNearest] = log1pf(s);
    __threadfence_block();  // every lane reads the whole row its wavefront wrote
    float y0, y1;
    mlp_eval_bf16(s_f[w], lane, w1, w2, P, norm, s_x[w], y0, y1);
    // every lane holds y0, y1 (reduced across the wavefront), so every lane forms the same values
    const float np = expf(y1) * nprobe_scale;
    const int c0 = np >= (float)nprobe ? nprobe : max(min_nprobe, (int)ceilf(np));
    const int qpred = min(max(c0, 1), nprobe);
    const float b =
        bound_scale > 0.f ? fminf(expm1f(fmaxf(y0, 0.f)) * bound_scale, 3.402823466e38f) : 3.402823466e38f;
    int c = qpred;
    // enough vectors for a full result: a query whose first lists hold fewer than kmin vectors
    // could not fill its top k (and would be redone over every list); its count grows to the
    // shortest prefix of its probe list that holds kmin
    if (list_off) {
        int64_t have = 0;
        for (int j = 0; j < c; ++j) {
            const int l = probes[q * pc + j];
            have += list_off[l + 1] - list_off[l];
        }
        while (have < kmin && c < nprobe) {
            const int l = probes[q * pc + c];
            have += list_off[l + 1] - list_off[l];
            ++c;
        }
    }
    const int ccov = c;
    if (margin > 0.f && b < 3.402823466e38f && lrad != nullptr) {
        const float thr = b * margin * (1.f + slack);
        const float d0 = fmaxf(pdist[q * pc], 0.f);
        const float* c0 = cent + (int64_t)probes[q * pc] * d;
        for (int j = ccov; j < nprobe; ++j) {  // wave-uniform loop over the ranks past the prediction
            const int l = probes[q * pc + j];
            const float dl = fmaxf(pdist[q * pc + j], 0.f);
            const float* cl = cent + (int64_t)l * d;
            float e = 0.f;
            for (int i = lane; i < d; i += 64) {
                const float t = cl[i] - c0[i];
                e = __builtin_fmaf(t, t, e);
            }
            for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
            const float rb = sqrtf(dl) * (1.f - slack) - lrad[l] * (1.f + slack);
            const float num = (dl - d0) - 2.f * slack * (dl + d0);
            const float vb = e > 0.f ? num / (2.f * sqrtf(e) * (1.f + slack)) : 0.f;
            const float g = fmaxf(rb, vb);
            if (g <= 0.f || g * g <= thr) c = j + 1;
        }
    }
    // the grouping's (list, rank) pair counts of the ranks this query searches
    if (bucket_cnt)
        for (int j = lane; j < min(c, nranks); j += 64) atomicAdd(&bucket_cnt[(int64_t)probes[q * pc + j] * nranks + j], 1);
    if (lane == 0) {
        qn[q] = c;
        if (c > ccov && widened) atomicAdd(widened, 1ull);
        bound[q] = b;
        if (bound2) bound2[q] = b;
        if (bound3) bound3[q] = b;
        pred[q * 2] = y0;
        pred[q * 2 + 1] = y1;
    }
}

void launch_mlp_serve(const float* near, const float* pdist, const int* probes, int pc, const float* Qp, int d,
                      int64_t n,
                      const __bf16* w1, const __bf16* w2, const float* P, const float* norm, float nprobe_scale,
                      float bound_scale, int min_nprobe, int nprobe, const float* lrad, const float* cent,
                      float margin, float slack, const int64_t* list_off, int kmin, int* qn, float* bound,
                      float* pred, float* bound2, float* bound3, unsigned long long* widened, int* bucket_cnt,
                      int nranks, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_serve, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, near, pdist, probes, pc, Qp, d, n, w1,
                       w2, P, norm, nprobe_scale, bound_scale, min_nprobe, nprobe, lrad, cent, margin, slack, list_off,
                       kmin, qn, bound, pred, bound2, bound3, widened, bucket_cnt, nranks);
}

// Radius of every list: the largest |x - c| over its vectors, from their squared residual norms
// (one wavefront per list; 0 for an empty list).
__global__ __launch_bounds__(256) void k_list_radius(const float* __restrict__ rnorm, const int64_t* __restrict__ list_off,
                                                     int nlist, float* __restrict__ lrad) {
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (l >= nlist) return;
    const int lane = threadIdx.x & 63;
    float m = 0.f;
    for (int64_t i = list_off[l] + lane; i < list_off[l + 1]; i += 64) m = fmaxf(m, rnorm[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) lrad[l] = sqrtf(m);
}

void launch_list_radius(const float* rnorm, const int64_t* list_off, int nlist, float* lrad, hipStream_t s) {
    if (nlist <= 0) return;
    hipLaunchKernelGGL(k_list_radius, dim3((unsigned)((nlist + 3) / 4)), dim3(256), 0, s, rnorm, list_off, nlist, lrad);
}

}  // namespace hv
