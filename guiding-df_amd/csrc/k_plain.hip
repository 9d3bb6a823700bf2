// Plain FP32 per-batch path: the comparator the mixed-precision fused path is
// measured against.  It mirrors the reference's structure — one kernel per
// encoding, one per MLP layer, one per head, activations round-tripping
// through memory (feature-major, like tcnn's column-major matrices) — but
// computes everything in fp32 (fp32 activations and accumulation; weights are
// the fp16 EMA inference weights widened to fp32), with no MFMA.
#include "device.h"
#include "internal.h"

namespace dfg {

struct RowInP {
    const float* row;
    int base_dims;
    float extra;
    __device__ float operator()(int d) const { return d < base_dims ? row[d] : extra; }
};

__global__ void plain_encode_kernel(const NetSpec* __restrict__ dns, const _Float16* tab, const float* feat, int stride,
                                    const float* extra, int64_t n, int64_t ld, float* out) {
    const NetSpec& ns = *dns;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const float* fr = feat + row * stride;
    const int base = extra ? ns.in_dims - 1 : ns.in_dims;
    const float ex = extra ? extra[row] : 0.0f;
    const int ew = ns.enc_width;
    encode_row(ns, tab, RowInP{fr, base, ex}, [=](int c, float v) {
        if (c < ew) out[(size_t)c * ld + row] = v;
    });
}

void launch_plain_encode(const NetSpec* dns, const _Float16* tab, const float* feat, int stride, const float* extra,
                         int64_t n, int64_t ld, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(plain_encode_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, dns, tab, feat, stride,
                       extra, n, ld, out);
}

// Y[m][row] = act(sum_k W[m][k] X[k][row]); W staged in LDS as fp32, transposed
// so the M weights of one k are a contiguous broadcast read.
template <int M>
__global__ __launch_bounds__(256) void plain_layer_kernel(const _Float16* __restrict__ W, int K, const float* __restrict__ X,
                                                          float* __restrict__ Y, int relu, int64_t n, int64_t ld) {
    extern __shared__ float wl[];  // [K][M]
    for (int i = threadIdx.x; i < K * M; i += blockDim.x) {
        const int m = i / K, k = i % K;
        wl[k * M + m] = (float)W[(size_t)m * K + k];
    }
    __syncthreads();
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float xk = X[(size_t)k * ld + row];
        const float4* wk = (const float4*)(wl + k * M);
#pragma unroll
        for (int m4 = 0; m4 < M / 4; ++m4) {
            const float4 w = wk[m4];
            acc[4 * m4 + 0] += w.x * xk;
            acc[4 * m4 + 1] += w.y * xk;
            acc[4 * m4 + 2] += w.z * xk;
            acc[4 * m4 + 3] += w.w * xk;
        }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) Y[(size_t)m * ld + row] = relu ? fmaxf(acc[m], 0.0f) : acc[m];
}

void launch_plain_layer(const _Float16* W, int M, int K, int K_alloc, const float* X, float* Y, int relu, int64_t n,
                        int64_t ld, hipStream_t s) {
    (void)K_alloc;
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256));
    const size_t lds = (size_t)K * M * 4;
    if (M == 64) hipLaunchKernelGGL((plain_layer_kernel<64>), grid, dim3(256), lds, s, W, K, X, Y, relu, n, ld);
    else if (M == 32) hipLaunchKernelGGL((plain_layer_kernel<32>), grid, dim3(256), lds, s, W, K, X, Y, relu, n, ld);
    else hipLaunchKernelGGL((plain_layer_kernel<16>), grid, dim3(256), lds, s, W, K, X, Y, relu, n, ld);
}

// Head of one net (phase 0: theta, phase 1: phi) — thetaPostInference /
// phiPostInference (train.h:895-929) in fp32 on feature-major logits.
template <int N>
__global__ __launch_bounds__(64) void plain_head_kernel(int phase, int n_theta, int linear, const float* __restrict__ lg,
                                                        int64_t ld, int64_t n, const float* __restrict__ u,
                                                        const uint8_t* __restrict__ guided, float* __restrict__ wi,
                                                        float* __restrict__ theta_in) {
    __shared__ float bins[64][N + 1];
    const int lane = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + lane;
    if (row >= n) return;
    float z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = lg[(size_t)i * ld + row];
    softmax_times_n<N>(z);
    float* P = bins[lane];
#pragma unroll
    for (int i = 0; i < N; ++i) P[i] = z[i];
    auto Pf = [&](int k) { return P[k]; };
    const bool g = u != nullptr && (guided == nullptr || guided[row] != 0);
    float* w = wi + 4 * row;
    float x = w[phase];
    if (g) {
        const float uv = u[4 * row + phase], ue = u[4 * row + 2 + phase];
        if (linear) { const Segment sg = search_linear<N>(z, uv); x = sample_linear<N>(sg, uv, ue, Pf); }
        else { const Segment sg = search_nearest<N>(z, uv); x = sample_nearest<N>(sg, uv, ue); }
    }
    const float p = linear ? pdf_linear<N>(x, Pf) : pdf_nearest<N>(x, Pf);
    w[phase] = x;
    w[2 + phase] = w[2 + phase] * p;
    if (phase == 0) theta_in[row] = theta_input(x, n_theta, linear);
}

void launch_plain_head(int phase, int n_bins, int n_theta, int linear, const float* logits, int64_t ld, int64_t n,
                       const float* u, const uint8_t* guided, float* wi, float* theta_in, hipStream_t s) {
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 63) / 64));
    if (n_bins == 16) hipLaunchKernelGGL((plain_head_kernel<16>), grid, dim3(64), 0, s, phase, n_theta, linear, logits, ld, n, u, guided, wi, theta_in);
    else if (n_bins == 32) hipLaunchKernelGGL((plain_head_kernel<32>), grid, dim3(64), 0, s, phase, n_theta, linear, logits, ld, n, u, guided, wi, theta_in);
    else hipLaunchKernelGGL((plain_head_kernel<64>), grid, dim3(64), 0, s, phase, n_theta, linear, logits, ld, n, u, guided, wi, theta_in);
}

__global__ void plain_rgb_kernel(const float* __restrict__ o, int64_t ld, int64_t n, float* __restrict__ rgb) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    for (int c = 0; c < 3; ++c) rgb[3 * row + c] = o[(size_t)c * ld + row];
}

void launch_plain_rgb(const float* out, int64_t ld, int64_t n, float* rgb, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(plain_rgb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, ld, n, rgb);
}

// ---------------------------------------------------------------------------
// Plain FP32 training step: the same stages as the reference's trainStep
// (learnedDistribution.cpp:230-290) — encode, one kernel per layer forward,
// loss head, one kernel per layer backward, weight gradients, grid scatter —
// with fp32 activations, bins, gradients and accumulation throughout.
// ---------------------------------------------------------------------------

// phi-net theta input of a training record (train.cu:52-60)
__global__ void plain_theta_target_kernel(const float* __restrict__ target, int64_t n, int n_theta, int linear,
                                          float* __restrict__ out) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row < n) out[row] = theta_input(target[row * 9], n_theta, linear);
}

void launch_plain_theta_target(const float* target, int64_t n, int n_theta, int linear, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(plain_theta_target_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, target, n,
                       n_theta, linear, out);
}

// dX[m][row] = mask(m, row) * sum_k W[k][m_begin + m] dY[k][row]: the data
// gradient through a layer W [K][K_row] (rows = the layer's outputs), masked
// by the ReLU of the stored input activation (mask > 0) when mask is given.
template <int MT>
__global__ __launch_bounds__(256) void plain_layer_bwd_kernel(const _Float16* __restrict__ W, int K_row, int m_begin,
                                                              int M, int K, const float* __restrict__ dY,
                                                              const float* __restrict__ mask, float* __restrict__ dX,
                                                              int64_t n, int64_t ld) {
    extern __shared__ float wl[];  // [K][MT]
    for (int i = threadIdx.x; i < K * MT; i += blockDim.x) {
        const int k = i / MT, m = i % MT;
        wl[i] = m < M ? (float)W[(size_t)k * K_row + m_begin + m] : 0.0f;
    }
    __syncthreads();
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float acc[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m] = 0.0f;
    for (int k = 0; k < K; ++k) {
        const float gk = dY[(size_t)k * ld + row];
        const float4* wk = (const float4*)(wl + k * MT);
#pragma unroll
        for (int m4 = 0; m4 < MT / 4; ++m4) {
            const float4 w = wk[m4];
            acc[4 * m4 + 0] += w.x * gk;
            acc[4 * m4 + 1] += w.y * gk;
            acc[4 * m4 + 2] += w.z * gk;
            acc[4 * m4 + 3] += w.w * gk;
        }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        if (m < M) {
            const bool on = mask == nullptr || mask[(size_t)m * ld + row] > 0.0f;
            dX[(size_t)m * ld + row] = on ? acc[m] : 0.0f;
        }
    }
}

void launch_plain_layer_bwd(const _Float16* W, int K_row, int m_begin, int M, int K, const float* dY,
                            const float* mask, float* dX, int64_t n, int64_t ld, hipStream_t s) {
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (M <= 32)
        hipLaunchKernelGGL((plain_layer_bwd_kernel<32>), grid, dim3(256), (size_t)K * 32 * 4, s, W, K_row, m_begin,
                           M, K, dY, mask, dX, n, ld);
    else
        hipLaunchKernelGGL((plain_layer_bwd_kernel<64>), grid, dim3(256), (size_t)K * 64 * 4, s, W, K_row, m_begin,
                           M, K, dY, mask, dX, n, ld);
}

// DF loss head in fp32 (combine_AR_pdf_and_dl_doutput, train.h:931-965, with
// fp32 bins instead of the reference's fp16 storage); feature-major logits in,
// feature-major loss-scaled logit gradients out.
template <int NTH, int NPH>
__global__ void plain_df_loss_kernel(const float* __restrict__ target, const float* __restrict__ lt,
                                     const float* __restrict__ lp, int linear, float* __restrict__ g_th,
                                     float* __restrict__ g_ph, float* __restrict__ loss, float loss_scale, int64_t n,
                                     int64_t ld) {
#pragma clang fp contract(off)
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float zt[NTH], zp[NPH];
#pragma unroll
    for (int i = 0; i < NTH; ++i) zt[i] = lt[(size_t)i * ld + row];
#pragma unroll
    for (int i = 0; i < NPH; ++i) zp[i] = lp[(size_t)i * ld + row];
    softmax_times_n<NTH>(zt);
    softmax_times_n<NPH>(zp);
    const float* t = target + row * 9;
    auto Pt = [&](int k) { return zt[k]; };
    auto Pp = [&](int k) { return zp[k]; };
    const float x0 = t[0], x1 = t[1];
    const float pt = linear ? pdf_linear<NTH>(x0, Pt) : pdf_nearest<NTH>(x0, Pt);
    const float pp = linear ? pdf_linear<NPH>(x1, Pp) : pdf_nearest<NPH>(x1, Pp);
    const float apt = fabsf(pt), app = fabsf(pp);
    const float Li = (t[2] + t[3] + t[4]) / 3.0f;
    const float wiPdf = t[5] + DFG_EPS;
    const float guided = ((apt * app) * DFG_INV_4PI) * (1.0f / DFG_INV_4PI) + DFG_EPS;
    const float prefix = -Li / wiPdf / guided * (loss_scale / (float)n);
    loss[row] = -Li / wiPdf * det_logf(guided);
    {
        const float inv_res = 1.0f / (float)NTH;
        float gt[NTH];
        df_logit_grad<NTH>(x0, prefix * app * (float)NTH, linear, [&](int k) { return zt[k] * inv_res; }, gt);
#pragma unroll
        for (int i = 0; i < NTH; ++i) g_th[(size_t)i * ld + row] = loss_grad(gt[i]);
    }
    {
        const float inv_res = 1.0f / (float)NPH;
        float gp[NPH];
        df_logit_grad<NPH>(x1, prefix * apt * (float)NPH, linear, [&](int k) { return zp[k] * inv_res; }, gp);
#pragma unroll
        for (int i = 0; i < NPH; ++i) g_ph[(size_t)i * ld + row] = loss_grad(gp[i]);
    }
}

void launch_plain_df_loss(const float* target, const float* lt, const float* lp, int n_theta, int n_phi, int linear,
                          float* g_th, float* g_ph, float* loss, float loss_scale, int64_t n, int64_t ld,
                          hipStream_t s) {
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 127) / 128));
#define DFG_PLOSS(A, B)                                                                                       \
    if (n_theta == A && n_phi == B) {                                                                         \
        hipLaunchKernelGGL((plain_df_loss_kernel<A, B>), grid, dim3(128), 0, s, target, lt, lp, linear, g_th, \
                           g_ph, loss, loss_scale, n, ld);                                                     \
        return;                                                                                               \
    }
    DFG_PLOSS(16, 32) DFG_PLOSS(32, 32) DFG_PLOSS(64, 64) DFG_PLOSS(16, 16) DFG_PLOSS(32, 64)
#undef DFG_PLOSS
}

// relative_l2_luminance_loss (relative_l2_luminance.h:39-88) on fp32 predictions,
// feature-major [16][ld] in and out (outputs 3..15 are padding: zero gradient).
__global__ void plain_nrc_loss_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                      float* __restrict__ g, float* __restrict__ loss, float loss_scale, int64_t n,
                                      int64_t ld) {
#pragma clang fp contract(off)
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float p[3];
    for (int c = 0; c < 3; ++c) p[c] = pred[(size_t)c * ld + row];
    const float lum = 0.299f * p[0] + 0.587f * p[1] + 0.114f * p[2];
    const float den = lum * lum + 0.01f;
    const float n_total = (float)(n * 3);
    float lsum = 0.0f;
    for (int c = 0; c < 16; ++c) {
        float gv = 0.0f;
        if (c < 3) {
            const float diff = p[c] - target[row * 3 + c];
            lsum += diff * diff / den / 1.0f / n_total;
            gv = loss_grad(loss_scale * (2.0f * diff / den / 1.0f) / n_total);
        }
        g[(size_t)c * ld + row] = gv;
    }
    loss[row] = lsum;
}

void launch_plain_nrc_loss(const float* pred, const float* target, float* g, float* loss, float loss_scale, int64_t n,
                           int64_t ld, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(plain_nrc_loss_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pred, target, g,
                       loss, loss_scale, n, ld);
}

// Weight gradient of one layer over one chunk of rows: partial[c][m*K + k] =
// sum_{row in chunk, row < n} dY[m][row] X[k][row]; rows staged 32 at a time
// through LDS, one thread per (m, k) group, fp32 accumulation.  The chunk
// partials are reduced in chunk order afterwards.
constexpr int kPlainWgMaxM = 64;
constexpr int kPlainWgMaxK = DFG_MAX_ENC;
constexpr int kPlainWgPer = kPlainWgMaxM * kPlainWgMaxK / 256;

__global__ __launch_bounds__(256) void plain_wgrad_kernel(const float* __restrict__ dY, int M,
                                                          const float* __restrict__ X, int K,
                                                          float* __restrict__ partial, int64_t pstride, int64_t n,
                                                          int64_t ld, int chunk) {
    __shared__ float sg[kPlainWgMaxM][33];
    __shared__ float sx[kPlainWgMaxK][33];
    const int c = blockIdx.x;
    const int64_t r0 = (int64_t)c * chunk;
    int64_t r1 = r0 + chunk;
    if (r1 > n) r1 = n;
    const int t = threadIdx.x;
    const int total = M * K;
    float acc[kPlainWgPer];
#pragma unroll
    for (int j = 0; j < kPlainWgPer; ++j) acc[j] = 0.0f;
    for (int64_t rb = r0; rb < r1; rb += 32) {
        __syncthreads();
        for (int i = t; i < M * 32; i += 256) {
            const int m = i / 32, r = i % 32;
            sg[m][r] = rb + r < r1 ? dY[(size_t)m * ld + rb + r] : 0.0f;
        }
        for (int i = t; i < K * 32; i += 256) {
            const int k = i / 32, r = i % 32;
            sx[k][r] = rb + r < r1 ? X[(size_t)k * ld + rb + r] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPlainWgPer; ++j) {
            const int o = t + 256 * j;
            if (o < total) {
                const int m = o / K, k = o % K;
                float a = acc[j];
                for (int r = 0; r < 32; ++r) a += sg[m][r] * sx[k][r];
                acc[j] = a;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPlainWgPer; ++j) {
        const int o = t + 256 * j;
        if (o < total) partial[(size_t)c * pstride + o] = acc[j];
    }
}

void launch_plain_wgrad(const float* dY, int M, const float* X, int K, float* partial, int64_t pstride, int64_t n,
                        int64_t ld, int chunk, int n_chunks, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(plain_wgrad_kernel, dim3((unsigned)n_chunks), dim3(256), 0, s, dY, M, X, K, partial, pstride,
                       n, ld, chunk);
}

}  // namespace dfg
