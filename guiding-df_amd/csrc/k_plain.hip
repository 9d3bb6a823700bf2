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

}  // namespace dfg
