// Training step kernels: forward with stored activations, DF / NRC loss heads,
// MFMA backward (data gradients), MFMA weight gradients with a deterministic
// chunked reduction, grid scatter-add, fused Adam + EMA, fragment packing.
// Semantics: LearnedDistribution::trainStep (learnedDistribution.cpp:230-290),
// tcnn fully_fused_mlp.cu:754-854, grid.h:394-520, adam.h:47-120, ema.h:62-77.
#include "device.h"
#include "internal.h"

namespace dfg {

struct RowInT {
    const float* row;
    int base_dims;
    float extra;
    __device__ float operator()(int d) const { return d < base_dims ? row[d] : extra; }
};

// ---------------------------------------------------------------------------
// Forward of one net for a 64-row tile, storing the encoded input (x) and the
// hidden activations feature-major, and the output row-major as fp16
// (DF: softmax x N stored in half like softmaxMultFactor, train.h:874-893;
//  NRC: the raw network output).
// ---------------------------------------------------------------------------
template <int NT_OUT, int NBINS>
__global__ __launch_bounds__(64, 2) void train_fwd_kernel(const TrainFwdArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    // encoded input tile: fp16 (the values are rounded to fp16 anyway), row
    // stride xs halves; the fp32 output tile aliases it
    _Float16* xt = (_Float16*)smem;
    float* lo = (float*)smem;
    const NetSpec& ns = *a.ns;
    const int lane = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * 64;
    const int64_t row = row0 + lane;
    const int64_t r = row < a.n ? row : a.n - 1;
    const float* frow = a.feat + r * a.stride;
    float th_in = 0.0f;
    int base = ns.in_dims;
    if (a.target) {
        // generate_training_data_AR, train.cu:52-60: phi input theta from the record's direction
        th_in = theta_input(a.target[r * 9], a.theta_bins, a.linear);
        base = ns.in_dims - 1;
    }
    _Float16* const xrow = xt + lane * a.xs;
    // the training forward feeds fp16 activations, like the reference's
    encode_row(ns, a.tab, RowInT{frow, base, th_in}, [xrow](int c, float v) { xrow[c] = round_half(v); },
               a.gfeat ? a.gfeat + r : nullptr, a.ld);
    const int kc = ns.k_steps * 32;
    for (int c = 0; c < kc; ++c) a.x[(size_t)c * a.ld + row] = xrow[c];
    // the hidden activations are staged in the input tile (dead after layer 0)
    ActStore st{nullptr, a.act, a.ld, row0, (_Float16*)smem};
    mlp_forward_tile<NT_OUT, false>(ns, a.pack, xt, a.xs, lo, a.ls, &st);
    const float* lrow = lo + lane * a.ls;
    _Float16* o = a.out + row * (NT_OUT * 16);
    // row-major fp16 output, written 8 halves (16 bytes) at a time
    if (a.mode == 0) {
        float z[NBINS];
#pragma unroll
        for (int i = 0; i < NBINS; ++i) z[i] = lrow[i];
        softmax_times_n_half<NBINS>(z);
#pragma unroll
        for (int v = 0; v < NT_OUT * 2; ++v) {
            h8 hv;
#pragma unroll
            for (int j = 0; j < 8; ++j) hv[j] = v * 8 + j < NBINS ? (_Float16)z[v * 8 + j < NBINS ? v * 8 + j : 0] : (_Float16)0.0f;
            *(h8*)(o + v * 8) = hv;
        }
    } else {
#pragma unroll
        for (int v = 0; v < NT_OUT * 2; ++v) {
            h8 hv;
#pragma unroll
            for (int j = 0; j < 8; ++j) hv[j] = (_Float16)lrow[v * 8 + j];
            *(h8*)(o + v * 8) = hv;
        }
    }
}

void launch_train_fwd(const TrainFwdArgs& a, const NetSpec& hs, hipStream_t s) {
    if (a.n <= 0) return;
    const unsigned blocks = (unsigned)(a.ld / 64);
    size_t lds = (size_t)64 * a.xs * 2;
    const size_t l2 = (size_t)64 * a.ls * 4;
    if (l2 > lds) lds = l2;
    const size_t l3 = (size_t)64 * kTileS * 2;  // activation staging tile
    if (l3 > lds) lds = l3;
    const int nt = hs.n_out_pad / 16;
    if (a.mode == 1) {
        hipLaunchKernelGGL((train_fwd_kernel<1, 16>), dim3(blocks), dim3(64), lds, s, a);
        return;
    }
    if (nt == 1) hipLaunchKernelGGL((train_fwd_kernel<1, 16>), dim3(blocks), dim3(64), lds, s, a);
    else if (nt == 2) hipLaunchKernelGGL((train_fwd_kernel<2, 32>), dim3(blocks), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((train_fwd_kernel<4, 64>), dim3(blocks), dim3(64), lds, s, a);
}

// ---------------------------------------------------------------------------
// DF loss head: combine_AR_pdf_and_dl_doutput (train.h:931-965) with the
// hand-written softmax Jacobian of AR.h:50-104.
// ---------------------------------------------------------------------------
template <int N, int NPAD>
__device__ __forceinline__ void df_grad(float x, float pdf_other, float prefix, const _Float16* __restrict__ bins,
                                        int linear, _Float16* rm, _Float16* fm, int64_t ld, int64_t row) {
#pragma clang fp contract(off)
    const float inv_res = 1.0f / (float)N;
    const float multiplier = prefix * pdf_other * (float)N;
    float g[N];
    df_logit_grad<N>(x, multiplier, linear, [&](int k) { return (float)bins[k] * inv_res; }, g);
    // the row-major copy as 16-byte stores (8 halves), the feature-major one
    // element by element (consecutive rows across the lanes)
#pragma unroll
    for (int v = 0; v < NPAD / 8; ++v) {
        h8 hv;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = v * 8 + j;
            hv[j] = i < N ? sat_h(loss_grad(g[i < N ? i : 0])) : (_Float16)0.0f;
            fm[(size_t)i * ld + row] = hv[j];
        }
        *(h8*)(rm + row * NPAD + v * 8) = hv;
    }
}

template <int NTH, int NPH>
__global__ void df_loss_kernel(const DfLossArgs a) {
#pragma clang fp contract(off)
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.ld) return;
    if (row >= a.n) {
        for (int i = 0; i < NTH; i += 8) *(h8*)(a.gout_th_rm + row * NTH + i) = h8{};
        for (int i = 0; i < NPH; i += 8) *(h8*)(a.gout_ph_rm + row * NPH + i) = h8{};
        for (int i = 0; i < NTH; ++i) a.gout_th_fm[(size_t)i * a.ld + row] = (_Float16)0.0f;
        for (int i = 0; i < NPH; ++i) a.gout_ph_fm[(size_t)i * a.ld + row] = (_Float16)0.0f;
        return;
    }
    const float* t = a.target + row * 9;
    const _Float16* bt = a.bins_th + row * NTH;
    const _Float16* bp = a.bins_ph + row * NPH;
    auto Pt = [&](int k) { return (float)bt[k]; };
    auto Pp = [&](int k) { return (float)bp[k]; };
    const float x0 = t[0], x1 = t[1];
    const float pt = a.linear ? pdf_linear<NTH>(x0, Pt) : pdf_nearest<NTH>(x0, Pt);
    const float pp = a.linear ? pdf_linear<NPH>(x1, Pp) : pdf_nearest<NPH>(x1, Pp);
    // AutoRegressive(guidingData): mPdfTheta = |p_theta|, mPdfPhi = |p_phi| (AR.h:25-35)
    const float apt = fabsf(pt), app = fabsf(pp);
    const DfRowTerms rt = df_row_terms(t, apt, app, a.loss_scale / (float)a.n);
    a.loss[row] = rt.loss;
    df_grad<NTH, NTH>(x0, app, rt.prefix, bt, a.linear, a.gout_th_rm, a.gout_th_fm, a.ld, row);
    df_grad<NPH, NPH>(x1, apt, rt.prefix, bp, a.linear, a.gout_ph_rm, a.gout_ph_fm, a.ld, row);
}

void launch_df_loss(const DfLossArgs& a, int nt, int np, hipStream_t s) {
    if (a.n <= 0) return;
    const unsigned blocks = (unsigned)((a.ld + 255) / 256);
#define DFG_LOSS(A, B)                                                                    \
    if (nt == A && np == B) {                                                             \
        hipLaunchKernelGGL((df_loss_kernel<A, B>), dim3(blocks), dim3(256), 0, s, a);     \
        return;                                                                           \
    }
    DFG_LOSS(16, 32) DFG_LOSS(32, 32) DFG_LOSS(64, 64) DFG_LOSS(16, 16) DFG_LOSS(32, 64)
#undef DFG_LOSS
}

// relative_l2_luminance_loss, tcnn relative_l2_luminance.h:39-88 (stride 16, dims 3)
__global__ void nrc_loss_kernel(const NrcLossArgs a) {
#pragma clang fp contract(off)
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.ld) return;
    if (row >= a.n) {
        *(h8*)(a.gout_rm + row * 16) = h8{};
        *(h8*)(a.gout_rm + row * 16 + 8) = h8{};
        for (int c = 0; c < 16; ++c) a.gout_fm[(size_t)c * a.ld + row] = (_Float16)0.0f;
        return;
    }
    const _Float16* p = a.pred + row * 16;
    const float r = (float)p[0], g = (float)p[1], b = (float)p[2];
    const float lum = 0.299f * r + 0.587f * g + 0.114f * b;
    const float den = lum * lum + 0.01f;
    const float n_total = (float)(a.n * 3);
    float lsum = 0.0f;
    h8 g0 = h8{};
    for (int c = 0; c < 16; ++c) {
        _Float16 gh = (_Float16)0.0f;
        if (c < 3) {
            const float diff = (float)p[c] - a.target[row * 3 + c];
            lsum += diff * diff / den / 1.0f / n_total;
            const float grad = 2.0f * diff / den / 1.0f;
            gh = sat_h(loss_grad(a.loss_scale * grad / n_total));
            g0[c] = gh;
        }
        a.gout_fm[(size_t)c * a.ld + row] = gh;
    }
    *(h8*)(a.gout_rm + row * 16) = g0;
    *(h8*)(a.gout_rm + row * 16 + 8) = h8{};
    a.loss[row] = lsum;
}

void launch_nrc_loss(const NrcLossArgs& a, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(nrc_loss_kernel, dim3((unsigned)((a.ld + 255) / 256)), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Backward data gradients for a 64-row tile (kernel_mlp_fused_backward,
// fully_fused_mlp.cu:150-259, plus the last-layer and input-gradient GEMMs).
// ---------------------------------------------------------------------------
// ReLU mask of the data gradient by the stored activation (in the tile), and
// the masked gradient written back into the tile.
__device__ __forceinline__ void mask_tile(f4 (&c)[4], _Float16* tile, int q, int i16, int g) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            _Float16* e = tile + (16 * t + 4 * g + r) * kTileS + q * 16 + i16;
            const float v = (float)*e > 0.0f ? c[t][r] : 0.0f;
            c[t][r] = v;
            *e = sat_h(v);
        }
}

template <int KS_OUT>
__global__ __launch_bounds__(64) void train_bwd_kernel(const TrainBwdArgs a) {
    __shared__ __align__(16) _Float16 tile[64 * kTileS];
    const NetSpec& ns = *a.ns;
    const int lane = threadIdx.x;
    const int i16 = lane & 15, g = lane >> 4;
    const int64_t row0 = (int64_t)blockIdx.x * 64;
    const int H = ns.n_hidden;
    const int npad = ns.n_out_pad;
    const size_t lay = (size_t)64 * a.ld;
    f4 acc[4][4];
    h8 act[4][2];
    // layer tile: mask by a_l, write g_{l+1}, next operand
    auto mask_layer = [&](int l) {
        tile_load(tile, a.act + (size_t)l * lay, a.ld, row0, lane);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            mask_tile(acc[q], tile, q, i16, g);
            c_to_act(acc[q], act[q]);
        }
        __syncthreads();
        tile_store(tile, a.g + (size_t)l * lay, a.ld, row0, lane);
        __syncthreads();
    };
    // dL/da_H = W_H^T g_out (K = outputs, natural order)
    {
        const _Float16* pk = a.pack + ns.bwd_off[H];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KS_OUT; ++s) {
            h8 b[4];
            const int col = s * 32 + g * 8;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (col < npad) b[q] = *(const h8*)(a.gout + (row0 + q * 16 + i16) * npad + col);
                else b[q] = h8{0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const h8 w = ld_frag(pk, t, s, KS_OUT, lane);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q][t] = mfma16(w, b[q], acc[q][t]);
            }
        }
        mask_layer(H - 1);
    }
    // hidden layers
    for (int l = H - 1; l >= 1; --l) {
        const _Float16* pk = a.pack + ns.bwd_off[l];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const h8 w0 = ld_frag(pk, t, 0, 2, lane);
            const h8 w1 = ld_frag(pk, t, 1, 2, lane);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f4 c = mfma16(w0, act[q][0], f4{0.f, 0.f, 0.f, 0.f});
                acc[q][t] = mfma16(w1, act[q][1], c);
            }
        }
        mask_layer(l - 1);
    }
    // dL/dx = W_0^T g_1, only the requested input columns (the grid features),
    // 16 columns at a time through LDS ([column][row] fp32, row stride 68)
    if (a.dx) {
        float* ft = (float*)tile;
        const _Float16* pk = a.pack + ns.bwd_off[0];
        const int t_lo = a.dx_begin / 16;
        const int t_hi = (a.dx_begin + a.dx_n + 15) / 16;
        for (int t = t_lo; t < t_hi; ++t) {
            const h8 w0 = ld_frag(pk, t, 0, 2, lane);
            const h8 w1 = ld_frag(pk, t, 1, 2, lane);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                f4 c = mfma16(w0, act[q][0], f4{0.f, 0.f, 0.f, 0.f});
                c = mfma16(w1, act[q][1], c);
#pragma unroll
                for (int r = 0; r < 4; ++r) ft[(4 * g + r) * 68 + q * 16 + i16] = c[r];
            }
            __syncthreads();
            // 16 columns x 64 rows: 16 lanes per column, 4 rows each
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int cc = (lane >> 4) + 4 * k;
                const int col = 16 * t + cc - a.dx_begin;
                if (col >= 0 && col < a.dx_n)
                    *(float4*)(a.dx + (size_t)col * a.ld + row0 + 4 * i16) = *(const float4*)(ft + cc * 68 + 4 * i16);
            }
            __syncthreads();
        }
    }
}

void launch_train_bwd(const TrainBwdArgs& a, const NetSpec& hs, hipStream_t s) {
    if (a.n <= 0) return;
    const unsigned blocks = (unsigned)(a.ld / 64);
    const int ks = (hs.n_out_pad + 31) / 32;
    if (ks == 1) hipLaunchKernelGGL((train_bwd_kernel<1>), dim3(blocks), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((train_bwd_kernel<2>), dim3(blocks), dim3(64), 0, s, a);
}

// ---------------------------------------------------------------------------
// Weight gradients dW_l = G_{l+1} A_l^T summed over the batch, on MFMA with the
// batch as the K dimension.  Block (chunk c, layer l): each of the 4 waves
// takes a quarter of the chunk's K-steps and every output tile (all M rows, N
// columns 64 at a time), so each fragment is loaded by one wave only; the
// waves' sums are added in LDS in a fixed tree ((w0 + w2) + (w1 + w3)) and go
// to partial[c], which is reduced in chunk order: the result does not depend
// on scheduling.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 2) void wgrad_kernel(const WgradArgs a) {
    __shared__ __align__(16) float red[2][64 * 64];
    const NetSpec& ns = *a.ns;
    const int l = blockIdx.y;
    const int c = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g = lane >> 4;
    const int H = ns.n_hidden;
    const int M = l == H ? ns.n_out_pad : 64;
    const int N = l == 0 ? ns.enc_width : 64;
    const int mt = M / 16, nt_n = N / 16;
    const size_t lay = (size_t)64 * a.ld;
    const _Float16* G = l == H ? a.gout_fm : a.g + (size_t)l * lay;        // g_{l+1}
    const _Float16* A = l == 0 ? a.x : a.act + (size_t)(l - 1) * lay;      // a_l
    const int64_t k0 = (int64_t)c * a.chunk;
    int64_t k1 = k0 + a.chunk;
    if (k1 > a.ld) k1 = a.ld;
    const int nsteps = (int)((k1 - k0) / 32);
    const int s_lo = wave * nsteps / 4, s_hi = (wave + 1) * nsteps / 4;
    float* out = a.partial + (size_t)c * ns.n_matrix_params + ns.layer_off[l];
    for (int n0 = 0; n0 < nt_n; n0 += 4) {
        const int ntg = nt_n - n0 < 4 ? nt_n - n0 : 4;
        f4 acc[4][4];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[m][t] = f4{0.f, 0.f, 0.f, 0.f};
        // U K-steps at a time, all of their fragments loaded before the first product
        auto steps = [&](int s, auto U) {
            constexpr int NU = decltype(U)::value;
            h8 af[NU][4], bf[NU][4];
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int64_t kb = k0 + (int64_t)(s + u) * 32 + g * 8;
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    if (m < mt) af[u][m] = *(const h8*)(G + (size_t)(m * 16 + i16) * a.ld + kb);
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (t < ntg) bf[u][t] = *(const h8*)(A + (size_t)((n0 + t) * 16 + i16) * a.ld + kb);
            }
#pragma unroll
            for (int u = 0; u < NU; ++u)
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (m < mt && t < ntg) acc[m][t] = mfma16(af[u][m], bf[u][t], acc[m][t]);
        };
        int s = s_lo;
        for (; s + 2 <= s_hi; s += 2) steps(s, std::integral_constant<int, 2>{});
        if (s < s_hi) steps(s, std::integral_constant<int, 1>{});
        // acc[m][t][r]: row m*16 + 4g + r, column (n0 + t)*16 + i16 -> red[.][row*64 + t*16 + i16]
        auto put = [&](float* dst) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) dst[(m * 16 + 4 * g + r) * 64 + t * 16 + i16] = acc[m][t][r];
        };
        auto add = [&](const float* src) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[m][t][r] += src[(m * 16 + 4 * g + r) * 64 + t * 16 + i16];
        };
        if (wave >= 2) put(red[wave - 2]);
        __syncthreads();
        if (wave < 2) add(red[wave]);
        __syncthreads();
        if (wave == 1) put(red[0]);
        __syncthreads();
        if (wave == 0) {
            add(red[0]);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (m < mt && t < ntg) {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            out[(size_t)(m * 16 + 4 * g + r) * N + (n0 + t) * 16 + i16] = acc[m][t][r];
                    }
        }
        __syncthreads();
    }
}

// grad[i] = sum over chunks of partial[c][i], in a fixed order: block of 64
// params x 16 chunk groups, group g summing chunks g, g+16, ..., then the 16
// group sums added in order.
__global__ __launch_bounds__(1024) void wgrad_reduce_kernel(const float* __restrict__ partial, int n_chunks,
                                                            int64_t n_matrix, float* __restrict__ grad) {
    __shared__ float red[16][64];
    const int j = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + j;
    float s = 0.0f;
    if (i < n_matrix)
        for (int c = g; c < n_chunks; c += 16) s += partial[(size_t)c * n_matrix + i];
    red[g][j] = s;
    __syncthreads();
    if (g == 0 && i < n_matrix) {
        float t = red[0][j];
        for (int k = 1; k < 16; ++k) t += red[k][j];
        grad[i] = t;
    }
}

void launch_wgrad_reduce(const float* partial, int n_chunks, int64_t n_matrix, float* grad, hipStream_t s) {
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((n_matrix + 63) / 64)), dim3(1024), 0, s, partial,
                       n_chunks, n_matrix, grad);
}

void launch_wgrad(const WgradArgs& a, const NetSpec& hs, hipStream_t s) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(wgrad_kernel, dim3((unsigned)a.n_chunks, (unsigned)(hs.n_hidden + 1)), dim3(256), 0, s, a);
    launch_wgrad_reduce(a.partial, a.n_chunks, (int64_t)hs.n_matrix_params, a.grad, s);
}

// ---------------------------------------------------------------------------
// Grid backward: scatter w * dL/dy into the table gradient (kernel_grid_backward,
// grid.h:394-520), fp32 atomics.
// ---------------------------------------------------------------------------
__global__ void grid_bwd_kernel(const GridBwdArgs a, int n_groups) {
    const NetSpec& ns = *a.ns;
    const GridSpec& gs = ns.grid;
    // XCD-aware: block b scatters only into levels l % n_groups == b % n_groups,
    // so each XCD's atomics stay inside its own slice of the table
    const int grp = blockIdx.x % n_groups;
    const int64_t row = (int64_t)(blockIdx.x / n_groups) * blockDim.x + threadIdx.x;
    if (row >= a.n) return;
    const float* fr = a.feat + row * a.stride + a.pos_dim;
    const int F = gs.n_feat;
    for (int level = grp; level < gs.n_levels; level += n_groups) {
        float* gg = a.grad + gs.param_begin + (size_t)gs.offset[level] * F;
        float d[8];
        for (int f = 0; f < F; ++f) d[f] = a.dx[(size_t)(level * F + f) * a.ld + row];
        GridCorners c;
        grid_corners(gs, level, fr[0], fr[1], fr[2], c);
        if (gs.interp == INTERP_NEAREST) {
            for (int f = 0; f < F; ++f) unsafeAtomicAdd(gg + (size_t)c.idx[0] * F + f, d[f]);
            continue;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            for (int f = 0; f < F; ++f) unsafeAtomicAdd(gg + (size_t)c.idx[k] * F + f, c.w[k] * d[f]);
    }
}

void launch_grid_bwd(const GridBwdArgs& a, const NetSpec& hs, hipStream_t s) {
    if (a.n <= 0 || !hs.grid.present) return;
    const int groups = hs.grid.n_levels < kXcdGroups ? hs.grid.n_levels : kXcdGroups;
    hipLaunchKernelGGL(grid_bwd_kernel, dim3((unsigned)(((a.n + 255) / 256) * groups)), dim3(256), 0, s, a, groups);
}

// ---------------------------------------------------------------------------
// Bucketed grid backward.  Scattered fp32 atomics run at a fixed device-wide
// rate (~21 G/s on MI355X whatever their scope or XCD placement), so the
// scatter of kernel_grid_backward (grid.h:394-520) is restructured as a
// counting sort by table slice:
//   count   : per (row chunk, level) block, updates per slice of the level
//   scan    : exclusive offsets in (slice, chunk) order, the work list
//   scatter : records {entry in slice, w * dL/dy[f]} to their slice's range,
//             and the largest |update| of each level
//   reduce  : one block per (slice, part) sums its records in LDS and writes
//             the slice of the gradient once (every table entry is written, so
//             the gradient needs no zero fill); parts of a slice cut into
//             several parts write partial sums, which combine adds in order
// The LDS sums are 64-bit fixed point: fp32 LDS atomics are serialised per
// lane on gfx950 (~0.2 T updates/s device-wide) while 64-bit integer LDS
// atomics run at ~2.4 T/s.  A slice with R records and level maximum
// |update| < 2^e uses the scale 2^S, S = 62 - bits(R) - e, so no partial sum
// can overflow and every update keeps ~40 or more bits below the level's
// largest; integer sums do not depend on the order of the additions, so the
// grid gradient is deterministic.  A level whose maximum is not finite gets
// NaN gradients.
// ---------------------------------------------------------------------------
// rint(v * 2^S) as an integer, from v's bits: v = +-m 2^(E - 150) (m with
// its implicit bit; subnormals as E = 1 without it), so v 2^S = +-m 2^sh with
// sh = E - 150 + S.  sh >= 0 is a left shift (the scale bounds the result
// below 2^62); sh < 0 a right shift rounded half to even (0 for sh <= -25,
// where m 2^sh < 1/2).
__device__ __forceinline__ long long fix64(float v, int S) {
    const uint32_t b = __float_as_uint(v);
    const uint32_t E = (b >> 23) & 0xffu;
    const uint64_t m = (b & 0x7fffffu) | (E ? 0x800000u : 0u);
    const int sh = (int)(E ? E : 1u) - 150 + S;
    uint64_t q;
    if (sh >= 0) {
        q = sh < 64 ? m << sh : 0ull;  // >= 64 only for inf / NaN, whose level is flagged non-finite
    } else if (sh <= -25) {
        q = 0;
    } else {
        const int k = -sh;
        q = m >> k;
        const uint64_t r = m & ((1ull << k) - 1ull), half = 1ull << (k - 1);
        q += (r > half || (r == half && (q & 1ull))) ? 1ull : 0ull;
    }
    return (b >> 31) ? -(long long)q : (long long)q;
}

// q 2^-S: the int64 -> float conversion rounds once to nearest, the scaling
// by a power of two is exact (above the subnormal range).
__device__ __forceinline__ float unfix64(unsigned long long q, int S) {
    return __builtin_ldexpf((float)(long long)q, -S);
}

// Scale exponent of slice s (level `level`); finite = false when the level's
// largest update is inf or NaN.
__device__ __forceinline__ int slice_scale(const GridBucketArgs& a, const uint32_t* lvl_max, int level, int s,
                                           bool& finite) {
    const float mx = __uint_as_float(lvl_max[level]);
    finite = mx <= 3.402823466e+38f;
    if (!finite || !(mx > 0.0f)) return 0;
    int e;
    (void)frexpf(mx, &e);  // mx < 2^e
    const int64_t R = a.offs[s + 1] - a.offs[s];
    const int lb = R > 0 ? 64 - __clzll((unsigned long long)R) : 0;  // R < 2^lb
    return 62 - lb - e;
}
__device__ __forceinline__ int level_slices(const GridSpec& g, int l, int shift) {
    const uint32_t e = g.offset[l + 1] - g.offset[l];
    return (int)((e + (1u << shift) - 1) >> shift);
}

__device__ __forceinline__ int level_slice_base(const GridSpec& g, int l, int shift) {
    int b = 0;
    for (int k = 0; k < l; ++k) b += level_slices(g, k, shift);
    return b;
}

// Records: F == 2 as one 12-byte {entry, g0, g1} store; other F as an
// entry array plus an [F] value array.
__device__ __forceinline__ void put_record(const GridBucketArgs& a, int F, int64_t pos, uint32_t e, float w,
                                           const float (&d)[8]) {
    if (F == 2) {
        ((uint3*)a.rec_val)[pos] = make_uint3(e, __float_as_uint(w * d[0]), __float_as_uint(w * d[1]));
    } else {
        a.rec_idx[pos] = e;
        for (int f = 0; f < F; ++f) a.rec_val[pos * F + f] = w * d[f];
    }
}

// Per (row chunk, level) block: updates per slice of the level, kept per
// (slice, chunk) and added to the slice totals.
__global__ __launch_bounds__(1024) void grid_bwd_count_kernel(const GridBucketArgs a) {
    __shared__ uint32_t hist[kGridMaxSlicesPerLevel];
    const GridSpec& gs = a.ns->grid;
    const int level = blockIdx.y;
    const int c = blockIdx.x;
    const int nsl = level_slices(gs, level, a.slice_shift);
    for (int i = threadIdx.x; i < nsl; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const int64_t r0 = (int64_t)c * kGridChunkRows;
    const int64_t r1 = r0 + kGridChunkRows < a.n ? r0 + kGridChunkRows : a.n;
    const int nc = gs.interp == INTERP_NEAREST ? 1 : 8;
    for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
        const float* fr = a.feat + row * a.stride + a.pos_dim;
        GridCorners cr;
        grid_corners(gs, level, fr[0], fr[1], fr[2], cr);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nc) atomicAdd(&hist[cr.idx[k] >> a.slice_shift], 1u);
    }
    __syncthreads();
    const int sb = level_slice_base(gs, level, a.slice_shift);
    for (int i = threadIdx.x; i < nsl; i += blockDim.x) {
        const uint32_t h = hist[i];
        a.cnt[(size_t)(sb + i) * a.n_chunks + c] = h;
        if (h) atomicAdd((unsigned long long*)&a.offs[sb + i], (unsigned long long)h);
    }
}

// Inclusive scan of one int64 per thread over a 1024-thread block.
__device__ __forceinline__ int64_t block_scan_1024(int64_t* part, int64_t x) {
    const int t = threadIdx.x;
    part[t] = x;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int64_t y = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    return part[t];
}

// One block: slice totals -> exclusive offsets (total to offs[n_slices]) and
// write cursors, then the reduce work list: each slice cut into parts of at
// most kGridPartRecords records.
__global__ __launch_bounds__(1024) void grid_bwd_scan_kernel(const GridBucketArgs a, int n_slices) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (n_slices + 1023) / 1024;
    int b = t * per;
    if (b > n_slices) b = n_slices;
    int e = b + per;
    if (e > n_slices) e = n_slices;
    int64_t sum = 0, np = 0, nm = 0, ns = 0;
    for (int i = b; i < e; ++i) {
        const int64_t c = a.offs[i];
        sum += c;
        const int64_t parts = c > 0 ? (c + kGridPartRecords - 1) / kGridPartRecords : 1;
        np += parts;
        if (parts > 1) {
            nm += parts;
            ns += 1;
        }
    }
    const int64_t inc = block_scan_1024(part, sum);
    __syncthreads();
    const int64_t inc_p = block_scan_1024(part, np);
    __syncthreads();
    const int64_t inc_m = block_scan_1024(part, nm);
    __syncthreads();
    const int64_t inc_s = block_scan_1024(part, ns);
    int64_t run = inc - sum, it = inc_p - np, slot = inc_m - nm, ms = inc_s - ns;
    for (int i = b; i < e; ++i) {
        const int64_t c = a.offs[i];
        a.offs[i] = run;
        a.cursor[i] = run;
        run += c;
        const int parts = c > 0 ? (int)((c + kGridPartRecords - 1) / kGridPartRecords) : 1;
        for (int p = 0; p < parts; ++p) a.items[it++] = make_int4(i, p, parts, parts > 1 ? (int)(slot + p) : -1);
        if (parts > 1) {
            a.multi[ms++] = make_int4(i, (int)slot, parts, 0);
            slot += parts;
        }
    }
    if (t == 1023) {
        a.offs[n_slices] = inc;
        a.n_items[0] = (int)inc_p;
        a.n_items[1] = (int)inc_s;
    }
}

// Per (row chunk, level) block: reserve this block's range of every slice,
// then write one record per (row, corner).
__global__ __launch_bounds__(1024) void grid_bwd_scatter_kernel(const GridBucketArgs a) {
    __shared__ uint32_t fill[kGridMaxSlicesPerLevel];
    __shared__ int64_t base[kGridMaxSlicesPerLevel];
    __shared__ uint32_t bmax;
    const GridSpec& gs = a.ns->grid;
    const int F = gs.n_feat;
    const int level = blockIdx.y;
    const int c = blockIdx.x;
    const int nsl = level_slices(gs, level, a.slice_shift);
    const int sb = level_slice_base(gs, level, a.slice_shift);
    if (threadIdx.x == 0) bmax = 0u;
    for (int i = threadIdx.x; i < nsl; i += blockDim.x) {
        fill[i] = 0u;
        const uint32_t h = a.cnt[(size_t)(sb + i) * a.n_chunks + c];
        base[i] = h ? (int64_t)atomicAdd((unsigned long long*)&a.cursor[sb + i], (unsigned long long)h) : 0;
    }
    __syncthreads();
    // largest |update| (bits of |v|: monotone for non-negative floats, NaN above inf)
    uint32_t umax = 0u;
    const int64_t r0 = (int64_t)c * kGridChunkRows;
    const int64_t r1 = r0 + kGridChunkRows < a.n ? r0 + kGridChunkRows : a.n;
    const uint32_t emask = (1u << a.slice_shift) - 1u;
    for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
        float d[8];
        for (int f = 0; f < F; ++f) d[f] = a.dx[(size_t)(level * F + f) * a.ld + row];
        const float* fr = a.feat + row * a.stride + a.pos_dim;
        GridCorners cr;
        grid_corners(gs, level, fr[0], fr[1], fr[2], cr);
        if (gs.interp == INTERP_NEAREST) {
            const uint32_t sl = cr.idx[0] >> a.slice_shift;
            const int64_t pos = base[sl] + atomicAdd(&fill[sl], 1u);
            put_record(a, F, pos, cr.idx[0] & emask, 1.0f, d);
            for (int f = 0; f < F; ++f) umax = max(umax, __float_as_uint(fabsf(d[f])));
            continue;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t sl = cr.idx[k] >> a.slice_shift;
            const int64_t pos = base[sl] + atomicAdd(&fill[sl], 1u);
            put_record(a, F, pos, cr.idx[k] & emask, cr.w[k], d);
            for (int f = 0; f < F; ++f) umax = max(umax, __float_as_uint(fabsf(cr.w[k] * d[f])));
        }
    }
    atomicMax(&bmax, umax);
    __syncthreads();
    if (threadIdx.x == 0 && bmax) atomicMax(&a.lvl_max[level], bmax);
}

// Staged scatter (F = 2, 8 corners): the records of 448 rows at a time (four
// threads per row, one x pair of corners each; threads 0..767 take a second
// row) are sorted by slice in LDS and then written as one contiguous run per
// slice, instead of one scattered 12-byte store per record.  Longer runs
// leave fewer partly written lines: 256-row sub-chunks took 308 us at 262k
// samples, 128-row ones 408 us, 448-row ones 270 us.  Staged records are 20
// bytes (float4 values + entry | slice << 16), the slice cursors sit in
// dynamic LDS sized for the largest level, and 64 VGPRs keep two 1,024-thread
// blocks (75 KB of LDS each) per CU.  The same records land in the same
// slice ranges (only their order inside a range differs, which the
// order-independent fixed-point reduce does not see).
//
// PAIR: the theta and phi grids of identical structure in one pass (same
// positions, so the same corners, slices and record positions): one record
// per corner, {w dL/dy[0], w dL/dy[1], w dL/dy2[0], w dL/dy2[1]} as a float4
// and the 16-bit entry, stored in groups of 8 records (the 8 float4, then the
// 8 entries: 144 contiguous bytes, so a slice run is one contiguous store
// range rather than a value run plus a separate 2-byte-per-record entry run),
// and each net's level maxima.
constexpr int kStageRows = 448;  // rows per sub-chunk
template <bool PAIR>
__global__ __launch_bounds__(1024, 8) void grid_bwd_scatter_staged_kernel(const GridBucketArgs a) {
    // next record of each slice in this block's range (records < 2^31), the
    // sub-chunk's count and run offset of each slice: 3 x the level's slices
    extern __shared__ uint32_t slice_lds[];
    __shared__ float4 sval[kStageRows * 8];   // staged record values
    __shared__ uint32_t skey[kStageRows * 8];  // staged entry | slice << 16
    __shared__ uint32_t bmax, bmax2;
    const GridSpec& gs = a.ns->grid;
    const int level = blockIdx.y;
    const int c = blockIdx.x;
    const int t = threadIdx.x;
    const int nsl = level_slices(gs, level, a.slice_shift);
    const int sb = level_slice_base(gs, level, a.slice_shift);
    uint32_t* const next = slice_lds;
    uint32_t* const lcnt = slice_lds + nsl;
    uint32_t* const loff = slice_lds + 2 * nsl;
    if (t == 0) bmax = bmax2 = 0u;
    for (int i = t; i < nsl; i += blockDim.x) {
        const uint32_t h = a.cnt[(size_t)(sb + i) * a.n_chunks + c];
        next[i] = h ? (uint32_t)atomicAdd((unsigned long long*)&a.cursor[sb + i], (unsigned long long)h) : 0u;
    }
    const int64_t r0 = (int64_t)c * kGridChunkRows;
    const int64_t r1 = r0 + kGridChunkRows < a.n ? r0 + kGridChunkRows : a.n;
    const uint32_t emask = (1u << a.slice_shift) - 1u;
    uint32_t umax = 0u, umax2 = 0u;
    struct RowIn { float d0, d1, e0, e1, px, py, pz; };
    auto load_row = [&](int64_t row) {
        RowIn v{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (row < r1) {
            v.d0 = a.dx[(size_t)(level * 2) * a.ld + row];
            v.d1 = a.dx[(size_t)(level * 2 + 1) * a.ld + row];
            if constexpr (PAIR) {
                v.e0 = a.dx2[(size_t)(level * 2) * a.ld + row];
                v.e1 = a.dx2[(size_t)(level * 2 + 1) * a.ld + row];
            }
            const float* fr = a.feat + row * a.stride + a.pos_dim;
            v.px = fr[0];
            v.py = fr[1];
            v.pz = fr[2];
        }
        return v;
    };
    // this thread's rows of a sub-chunk: s0 + t/4 and (threads below
    // 4 (kStageRows - 256)) s0 + 256 + t/4; each thread one x pair of corners
    const int rq = t >> 2;
    const bool two = rq + 256 < kStageRows;
    RowIn cur[2];
    cur[0] = load_row(r0 + rq);
    cur[1] = two ? load_row(r0 + 256 + rq) : RowIn{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int64_t s0 = r0; s0 < r1; s0 += kStageRows) {
        for (int i = t; i < nsl; i += blockDim.x) lcnt[i] = 0u;
        __syncthreads();
        bool live[2];
        uint32_t sl[4] = {0u, 0u, 0u, 0u}, lp[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            live[h] = (h == 0 || two) && s0 + 256 * h + rq < r1;
            if (live[h]) {
                uint32_t idx[2];
                float w[2];
                grid_corner_pair(gs, level, cur[h].px, cur[h].py, cur[h].pz, t & 3, idx, w);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    sl[2 * h + j] = idx[j] >> a.slice_shift;
                    lp[2 * h + j] = atomicAdd(&lcnt[sl[2 * h + j]], 1u);
                }
            }
        }
        __syncthreads();
        // exclusive offsets of this sub-chunk's slice runs, by one wave
        if (t < 64) {
            const int per = (nsl + 63) / 64;
            const int b = t * per < nsl ? t * per : nsl;
            const int e = b + per < nsl ? b + per : nsl;
            uint32_t sum = 0u;
            for (int i = b; i < e; ++i) sum += lcnt[i];
            uint32_t inc = sum;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(inc, d, 64);
                if (t >= d) inc += y;
            }
            uint32_t run = inc - sum;
            for (int i = b; i < e; ++i) {
                loff[i] = run;
                run += lcnt[i];
            }
        }
        __syncthreads();
        // the records into their slice runs (corners recomputed: the same
        // operations give the same indices and weights, and nothing stays
        // live across the scan)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (live[h]) {
                uint32_t idx[2];
                float w[2];
                grid_corner_pair(gs, level, cur[h].px, cur[h].py, cur[h].pz, t & 3, idx, w);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const uint32_t k = loff[sl[2 * h + j]] + lp[2 * h + j];
                    const float v0 = w[j] * cur[h].d0, v1 = w[j] * cur[h].d1;
                    umax = max(umax, max(__float_as_uint(fabsf(v0)), __float_as_uint(fabsf(v1))));
                    float u0 = 0.0f, u1 = 0.0f;
                    if constexpr (PAIR) {
                        u0 = w[j] * cur[h].e0;
                        u1 = w[j] * cur[h].e1;
                        umax2 = max(umax2, max(__float_as_uint(fabsf(u0)), __float_as_uint(fabsf(u1))));
                    }
                    sval[k] = make_float4(v0, v1, u0, u1);
                    skey[k] = (idx[j] & emask) | (sl[2 * h + j] << 16);
                }
            }
        }
        __syncthreads();
        cur[0] = load_row(s0 + kStageRows + rq);
        if (two) cur[1] = load_row(s0 + kStageRows + 256 + rq);
        const int nrec = (int)((r1 - s0 < kStageRows ? r1 - s0 : kStageRows) * 8);
        for (int i = t; i < nrec; i += blockDim.x) {
            const uint32_t key = skey[i];
            const uint32_t s = key >> 16;
            const int64_t pos = (int64_t)next[s] + (int64_t)(i - (int)loff[s]);
            const float4 v = sval[i];
            if constexpr (PAIR) {
                unsigned char* grp = (unsigned char*)a.rec_val + (pos >> 3) * kPairGroupBytes;
                *(float4*)(grp + 16 * (pos & 7)) = v;
                *(uint16_t*)(grp + 128 + 2 * (pos & 7)) = (uint16_t)(key & 0xffffu);  // entry in a slice of kGridSliceFloats / 4
            } else {
                ((uint3*)a.rec_val)[pos] = make_uint3(key & 0xffffu, __float_as_uint(v.x), __float_as_uint(v.y));
            }
        }
        __syncthreads();
        for (int i = t; i < nsl; i += blockDim.x) next[i] += lcnt[i];
    }
    atomicMax(&bmax, umax);
    if constexpr (PAIR) atomicMax(&bmax2, umax2);
    __syncthreads();
    if (t == 0 && bmax) atomicMax(&a.lvl_max[level], bmax);
    if constexpr (PAIR) {
        if (t == 0 && bmax2) atomicMax(&a.lvl_max2[level], bmax2);
    }
}

// Level, first entry and value count of slice s.
__device__ __forceinline__ void slice_extent(const GridBucketArgs& a, int s, int& level, uint32_t& e0, int& nf) {
    const GridSpec& gs = a.ns->grid;
    int sb = 0;
    for (level = 0; level < gs.n_levels; ++level) {
        const int k = level_slices(gs, level, a.slice_shift);
        if (s < sb + k) break;
        sb += k;
    }
    const uint32_t lvl_entries = gs.offset[level + 1] - gs.offset[level];
    e0 = (uint32_t)(s - sb) << a.slice_shift;
    const uint32_t cap = 1u << a.slice_shift;
    const uint32_t ne = lvl_entries - e0 < cap ? lvl_entries - e0 : cap;
    nf = (int)ne * gs.n_feat;
}

// One block per (slice, part): sum the part's records into the slice in LDS
// as fixed-point integers (8 records in flight per thread), then write the
// slice (single part) or the part's partial sums (several parts).
__global__ __launch_bounds__(1024) void grid_bwd_reduce_kernel(const GridBucketArgs a) {
    // [F][NE]: feature-major, so random entries of a wave spread over the banks
    extern __shared__ unsigned long long acc[];
    if ((int)blockIdx.x >= *a.n_items) return;
    const int4 item = a.items[blockIdx.x];
    const GridSpec& gs = a.ns->grid;
    const int F = gs.n_feat;
    const int NE = kGridSliceFloats / F;
    const int s = item.x;
    int level, nf;
    uint32_t e0;
    slice_extent(a, s, level, e0, nf);
    bool finite;
    const int S = slice_scale(a, a.lvl_max, level, s, finite);
    for (int i = threadIdx.x; i < kGridSliceFloats; i += blockDim.x) acc[i] = 0ull;
    __syncthreads();
    const int64_t se = a.offs[s + 1];
    const int64_t pb = a.offs[s] + (int64_t)item.y * kGridPartRecords;
    const int64_t pe = pb + kGridPartRecords < se ? pb + kGridPartRecords : se;
    constexpr int U = 8;
    if (finite) {
        if (F == 2) {
            const uint3* r3 = (const uint3*)a.rec_val;
            for (int64_t p0 = pb + threadIdx.x; p0 < pe; p0 += (int64_t)U * blockDim.x) {
                uint3 r[U];
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    const int64_t p = p0 + (int64_t)j * blockDim.x;
                    r[j] = p < pe ? r3[p] : make_uint3(0u, 0u, 0u);
                }
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    if (p0 + (int64_t)j * blockDim.x < pe) {
                        atomicAdd(&acc[r[j].x], (unsigned long long)fix64(__uint_as_float(r[j].y), S));
                        atomicAdd(&acc[NE + r[j].x], (unsigned long long)fix64(__uint_as_float(r[j].z), S));
                    }
                }
            }
        } else {
            for (int64_t p = pb + threadIdx.x; p < pe; p += blockDim.x) {
                const uint32_t e = a.rec_idx[p];
                for (int f = 0; f < F; ++f) atomicAdd(&acc[f * NE + e], (unsigned long long)fix64(a.rec_val[p * F + f], S));
            }
        }
    }
    __syncthreads();
    if (item.z == 1) {
        float* g = a.grad + gs.param_begin + ((size_t)gs.offset[level] + e0) * F;
        for (int i = threadIdx.x; i < nf; i += blockDim.x)
            g[i] = finite ? unfix64(acc[(i % F) * NE + i / F], S) : __builtin_nanf("");
    } else {
        unsigned long long* pp = a.partial + (size_t)item.w * kGridSliceFloats;
        for (int i = threadIdx.x; i < kGridSliceFloats; i += blockDim.x) pp[i] = acc[i];
    }
}

// Slices cut into several parts: block (k, y) adds, for multi-part slice k
// and the y-th stripe of its values, the parts' partial sums in part order
// and writes the gradient.
__global__ __launch_bounds__(256) void grid_bwd_combine_kernel(const GridBucketArgs a) {
    if ((int)blockIdx.x >= a.n_items[1]) return;
    const int4 m = a.multi[blockIdx.x];
    const GridSpec& gs = a.ns->grid;
    const int F = gs.n_feat;
    int level, nf;
    uint32_t e0;
    slice_extent(a, m.x, level, e0, nf);
    bool finite;
    const int S = slice_scale(a, a.lvl_max, level, m.x, finite);
    float* g = a.grad + gs.param_begin + ((size_t)gs.offset[level] + e0) * F;
    const unsigned long long* pp = a.partial + (size_t)m.y * kGridSliceFloats;
    const int NE = kGridSliceFloats / F;  // partial sums are [F][NE]
    const int stripe = (nf + gridDim.y - 1) / gridDim.y;
    const int i0 = blockIdx.y * stripe;
    const int i1 = min(nf, i0 + stripe);
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int j = (i % F) * NE + i / F;
        unsigned long long q0 = 0ull, q1 = 0ull, q2 = 0ull, q3 = 0ull;
        int p = 0;
        for (; p + 4 <= m.z; p += 4) {
            q0 += pp[(size_t)p * kGridSliceFloats + j];
            q1 += pp[(size_t)(p + 1) * kGridSliceFloats + j];
            q2 += pp[(size_t)(p + 2) * kGridSliceFloats + j];
            q3 += pp[(size_t)(p + 3) * kGridSliceFloats + j];
        }
        for (; p < m.z; ++p) q0 += pp[(size_t)p * kGridSliceFloats + j];
        g[i] = finite ? unfix64(q0 + q1 + q2 + q3, S) : __builtin_nanf("");
    }
}

// Adam + EMA of grid entry e (F = 2 parameters) of both nets from its gradient
// (defined with the optimizer below).
__device__ void grid_adam_pair_entry(const GridAdam& ga, int64_t e, float2 gt, float2 gp);

// Paired theta/phi reduce: one block per (slice, part) of the 18-byte records
// (float4 values and 16-bit entries, in groups of 8); LDS sums [entry][4] (theta f0, f1, phi f0,
// f1), each net at its own fixed-point scale; writes both nets' slices (single
// part) or the part's partial sums.  With ga.on a single-part slice goes
// straight into the optimizer instead of the gradient arrays.
__global__ __launch_bounds__(1024) void grid_bwd_reduce_pair_kernel(const GridBucketArgs a, const GridAdam ga) {
    extern __shared__ unsigned long long acc[];  // [4][NE] (value-major: no bank conflicts)
    constexpr int NE = kGridSliceFloats / 4;
    static_assert(NE <= 65536, "paired record entries are 16-bit");
    const unsigned char* recs = (const unsigned char*)a.rec_val;
    if ((int)blockIdx.x >= *a.n_items) return;
    const int4 item = a.items[blockIdx.x];
    const GridSpec& gs = a.ns->grid;
    const int s = item.x;
    int level, nf2;
    uint32_t e0;
    slice_extent(a, s, level, e0, nf2);
    const int ne = nf2 / 2, nv = ne * 4;
    bool fin_a, fin_b;
    const int Sa = slice_scale(a, a.lvl_max, level, s, fin_a);
    const int Sb = slice_scale(a, a.lvl_max2, level, s, fin_b);
    (void)nv;
    for (int i = threadIdx.x; i < kGridSliceFloats; i += blockDim.x) acc[i] = 0ull;
    __syncthreads();
    const int64_t se = a.offs[s + 1];
    const int64_t pb = a.offs[s] + (int64_t)item.y * kGridPartRecords;
    const int64_t pe = pb + kGridPartRecords < se ? pb + kGridPartRecords : se;
    constexpr int U = 4;
    for (int64_t p0 = pb + threadIdx.x; p0 < pe; p0 += (int64_t)U * blockDim.x) {
        uint32_t e[U];
        float4 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int64_t p = p0 + (int64_t)j * blockDim.x;
            const bool in = p < pe;
            const unsigned char* grp = recs + (p >> 3) * kPairGroupBytes;
            e[j] = in ? (uint32_t) * (const uint16_t*)(grp + 128 + 2 * (p & 7)) : 0u;
            v[j] = in ? *(const float4*)(grp + 16 * (p & 7)) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            if (p0 + (int64_t)j * blockDim.x < pe) {
                unsigned long long* q = acc + e[j];
                if (fin_a) {
                    atomicAdd(q, (unsigned long long)fix64(v[j].x, Sa));
                    atomicAdd(q + NE, (unsigned long long)fix64(v[j].y, Sa));
                }
                if (fin_b) {
                    atomicAdd(q + 2 * NE, (unsigned long long)fix64(v[j].z, Sb));
                    atomicAdd(q + 3 * NE, (unsigned long long)fix64(v[j].w, Sb));
                }
            }
        }
    }
    __syncthreads();
    if (item.z == 1 && ga.on) {
        const int64_t E0 = (int64_t)gs.offset[level] + e0;
        const float nan = __builtin_nanf("");
        for (int en = threadIdx.x; en < ne; en += blockDim.x) {
            const float2 gt = fin_a ? make_float2(unfix64(acc[en], Sa), unfix64(acc[NE + en], Sa)) : make_float2(nan, nan);
            const float2 gp = fin_b ? make_float2(unfix64(acc[2 * NE + en], Sb), unfix64(acc[3 * NE + en], Sb))
                                    : make_float2(nan, nan);
            grid_adam_pair_entry(ga, E0 + en, gt, gp);
        }
    } else if (item.z == 1) {
        float* ga = a.grad + gs.param_begin + ((size_t)gs.offset[level] + e0) * 2;
        float* gb = a.grad2 + a.param_begin2 + ((size_t)gs.offset[level] + e0) * 2;
        for (int i = threadIdx.x; i < nf2; i += blockDim.x) {
            const int en = i >> 1, f = i & 1;
            ga[i] = fin_a ? unfix64(acc[f * NE + en], Sa) : __builtin_nanf("");
            gb[i] = fin_b ? unfix64(acc[(2 + f) * NE + en], Sb) : __builtin_nanf("");
        }
    } else {
        unsigned long long* pp = a.partial + (size_t)item.w * kGridSliceFloats;
        for (int i = threadIdx.x; i < kGridSliceFloats; i += blockDim.x) pp[i] = acc[i];
    }
}

// Multi-part slices of the paired records: block (k, y) adds, for slice k and
// the y-th stripe of its values, the parts' partial sums in part order (with
// ga.on: the y-th stripe of its entries, into the optimizer).
__global__ __launch_bounds__(256) void grid_bwd_combine_pair_kernel(const GridBucketArgs a, const GridAdam ga) {
    if ((int)blockIdx.x >= a.n_items[1]) return;
    const int4 m = a.multi[blockIdx.x];
    const GridSpec& gs = a.ns->grid;
    int level, nf2;
    uint32_t e0;
    slice_extent(a, m.x, level, e0, nf2);
    constexpr int NE = kGridSliceFloats / 4;  // partial sums are [4][NE]
    const int nv = nf2 * 2;  // output values: theta's nf2, then phi's nf2
    bool fin_a, fin_b;
    const int Sa = slice_scale(a, a.lvl_max, level, m.x, fin_a);
    const int Sb = slice_scale(a, a.lvl_max2, level, m.x, fin_b);
    const unsigned long long* pp = a.partial + (size_t)m.y * kGridSliceFloats;
    if (ga.on) {
        // (integer sums: any part order gives the same value)
        const int ne = nf2 / 2;
        const int estripe = (ne + gridDim.y - 1) / gridDim.y;
        const int en0 = blockIdx.y * estripe;
        const int en1 = min(ne, en0 + estripe);
        const int64_t E0 = (int64_t)gs.offset[level] + e0;
        const float nan = __builtin_nanf("");
        for (int en = en0 + threadIdx.x; en < en1; en += blockDim.x) {
            unsigned long long q[4] = {0ull, 0ull, 0ull, 0ull};
            for (int p = 0; p < m.z; ++p) {
                const unsigned long long* pr = pp + (size_t)p * kGridSliceFloats + en;
#pragma unroll
                for (int r = 0; r < 4; ++r) q[r] += pr[r * NE];
            }
            const float2 gt = fin_a ? make_float2(unfix64(q[0], Sa), unfix64(q[1], Sa)) : make_float2(nan, nan);
            const float2 gp = fin_b ? make_float2(unfix64(q[2], Sb), unfix64(q[3], Sb)) : make_float2(nan, nan);
            grid_adam_pair_entry(ga, E0 + en, gt, gp);
        }
        return;
    }
    float* ga_ = a.grad + gs.param_begin + ((size_t)gs.offset[level] + e0) * 2;
    float* gb_ = a.grad2 + a.param_begin2 + ((size_t)gs.offset[level] + e0) * 2;
    const int stripe = (nv + gridDim.y - 1) / gridDim.y;
    const int i0 = blockIdx.y * stripe;
    const int i1 = min(nv, i0 + stripe);
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const bool second = i >= nf2;
        const int o = second ? i - nf2 : i;  // output index in the net's slice: 2 en + f
        const int j = ((second ? 2 : 0) + (o & 1)) * NE + (o >> 1);
        unsigned long long q0 = 0ull, q1 = 0ull, q2 = 0ull, q3 = 0ull;
        int p = 0;
        for (; p + 4 <= m.z; p += 4) {
            q0 += pp[(size_t)p * kGridSliceFloats + j];
            q1 += pp[(size_t)(p + 1) * kGridSliceFloats + j];
            q2 += pp[(size_t)(p + 2) * kGridSliceFloats + j];
            q3 += pp[(size_t)(p + 3) * kGridSliceFloats + j];
        }
        for (; p < m.z; ++p) q0 += pp[(size_t)p * kGridSliceFloats + j];
        const unsigned long long q = q0 + q1 + q2 + q3;
        if (!second) ga_[o] = fin_a ? unfix64(q, Sa) : __builtin_nanf("");
        else gb_[o] = fin_b ? unfix64(q, Sb) : __builtin_nanf("");
    }
}

GridBucketPlan grid_bucket_plan(const NetSpec& hs, int64_t n, int values_per_entry) {
    GridBucketPlan p{};
    const GridSpec& g = hs.grid;
    const int F = g.n_feat;
    p.ok = hs.grid.present && (F == 1 || F == 2 || F == 4 || F == 8);
    if (!p.ok) return p;
    const int V = values_per_entry > 0 ? values_per_entry : F;
    int shift = 0;
    while ((1 << (shift + 1)) * V <= kGridSliceFloats) ++shift;
    p.slice_shift = shift;
    int maxl = 0;
    for (int l = 0; l < g.n_levels; ++l) {
        const uint32_t e = g.offset[l + 1] - g.offset[l];
        const int k = (int)((e + (1u << shift) - 1) >> shift);
        p.n_slices += k;
        if (k > maxl) maxl = k;
    }
    p.n_chunks = (int)((n + kGridChunkRows - 1) / kGridChunkRows);
    p.records = n * (int64_t)g.n_levels * (g.interp == INTERP_NEAREST ? 1 : 8);
    p.max_items = p.n_slices + (p.records + kGridPartRecords - 1) / kGridPartRecords;
    // a slice of c > P records has ceil(c / P) < 2c / P parts
    p.max_partials = 2 * (p.records / kGridPartRecords) + 1;
    p.max_multi = p.records / kGridPartRecords + 1;  // each holds more than P records
    p.max_level_slices = maxl;
    p.ok = maxl <= kGridMaxSlicesPerLevel && p.records <= ((int64_t)1 << 31) && p.n_chunks <= 65535 &&
           p.max_items < ((int64_t)1 << 30);
    return p;
}

void launch_grid_bwd_bucketed(const GridBucketArgs& a, const NetSpec& hs, const GridBucketPlan& p, hipStream_t s) {
    if (a.n <= 0 || !p.ok) return;
    const dim3 g1((unsigned)p.n_chunks, (unsigned)hs.grid.n_levels);
    // int64 slice totals, then the per-level maxima right after them
    launch_fill_u32((unsigned int*)a.offs, 0u, 2 * (int64_t)(p.n_slices + 1) + DFG_MAX_LEVELS, s);
    hipLaunchKernelGGL(grid_bwd_count_kernel, g1, dim3(1024), 0, s, a);
    hipLaunchKernelGGL(grid_bwd_scan_kernel, dim3(1), dim3(1024), 0, s, a, p.n_slices);
    if (hs.grid.n_feat == 2 && hs.grid.interp != INTERP_NEAREST)
        hipLaunchKernelGGL(grid_bwd_scatter_staged_kernel<false>, g1, dim3(1024),
                           (size_t)3 * p.max_level_slices * sizeof(uint32_t), s, a);
    else
        hipLaunchKernelGGL(grid_bwd_scatter_kernel, g1, dim3(1024), 0, s, a);
    const size_t lds = (size_t)kGridSliceFloats * sizeof(unsigned long long);
    hipLaunchKernelGGL(grid_bwd_reduce_kernel, dim3((unsigned)p.max_items), dim3(1024), lds, s, a);
    hipLaunchKernelGGL(grid_bwd_combine_kernel, dim3((unsigned)p.max_multi, 16), dim3(256), 0, s, a);
}

// Theta and phi grids of identical structure (F = 2, not nearest) in one
// bucketed pass: count, scan, the paired staged scatter (one 18-byte record
// per corner for both nets), then one reduce and combine writing both
// gradients.
void launch_grid_bwd_bucketed_pair(const GridBucketArgs& a, const NetSpec& hs, const GridBucketPlan& p,
                                   const GridAdam& ga, hipStream_t s) {
    if (a.n <= 0 || !p.ok) return;
    const dim3 g1((unsigned)p.n_chunks, (unsigned)hs.grid.n_levels);
    // int64 slice totals, then both nets' per-level maxima right after them
    launch_fill_u32((unsigned int*)a.offs, 0u, 2 * (int64_t)(p.n_slices + 1) + 2 * DFG_MAX_LEVELS, s);
    hipLaunchKernelGGL(grid_bwd_count_kernel, g1, dim3(1024), 0, s, a);
    hipLaunchKernelGGL(grid_bwd_scan_kernel, dim3(1), dim3(1024), 0, s, a, p.n_slices);
    hipLaunchKernelGGL(grid_bwd_scatter_staged_kernel<true>, g1, dim3(1024),
                       (size_t)3 * p.max_level_slices * sizeof(uint32_t), s, a);
    const size_t lds = (size_t)kGridSliceFloats * sizeof(unsigned long long);
    hipLaunchKernelGGL(grid_bwd_reduce_pair_kernel, dim3((unsigned)p.max_items), dim3(1024), lds, s, a, ga);
    hipLaunchKernelGGL(grid_bwd_combine_pair_kernel, dim3((unsigned)p.max_multi, 16), dim3(256), 0, s, a, ga);
}

// ---------------------------------------------------------------------------
// Adam (adam.h:47-120) fused with the half-precision EMA (ema.h:62-77).
// ---------------------------------------------------------------------------
// Adam + EMA of parameter i; returns its new fp16 weight and EMA.
__device__ __forceinline__ void adam_ema_param(const AdamArgs& a, int64_t i, _Float16& w16_out, _Float16& ema_out) {
#pragma clang fp contract(off)
    float gr = a.grad[i] / a.loss_scale;
    const bool matrix = i < a.n_matrix;
    _Float16 w16;
    if (matrix || gr != 0.0f) {
        const float wf = a.master[i];
        if (matrix) gr += a.l2_reg * wf;
        const float gsq = gr * gr;
        const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * gr;
        const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * gsq;
        a.m[i] = m;
        a.v[i] = v;
        uint32_t step;
        if (a.steps16) {
            step = min((uint32_t)a.steps16[i] + 1u, kShortStepMax);
            a.steps16[i] = (uint16_t)step;
        } else {
            step = a.steps[i] + 1;
            a.steps[i] = step;
        }
        // bias corrections beta^step as 2^(step log2 beta) (one hardware exp2
        // each instead of two powf per parameter)
        float lr = a.lr;
        lr *= sqrtf(1.0f - exp2f((float)step * a.log2_beta2)) / (1.0f - exp2f((float)step * a.log2_beta1));
        const float elr = fminf(fmaxf(lr / (sqrtf(v) + a.eps), 0.0f), 3.402823466e+38f);
        const float nw = wf - elr * m;
        a.master[i] = nw;
        w16 = round_half(nw);
        a.w16[i] = w16;
    } else {
        w16 = a.w16[i];
    }
    const float e = (float)a.ema16[i];
    const float w = (float)w16;
    const _Float16 ne = round_half((e * a.ema_decay * a.ema_debias_old + w * (1.0f - a.ema_decay)) * a.ema_debias_new);
    a.ema16[i] = ne;
    w16_out = w16;
    ema_out = ne;
}

__global__ void adam_ema_kernel(const AdamArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_params) return;
    _Float16 w16, ne;
    adam_ema_param(a, i, w16, ne);
    if (a.pair_w16 && i >= a.grid_begin) {
        // the interleaved theta/phi tables: entry {theta f0, f1, phi f0, f1}
        const int64_t l = i - a.grid_begin;
        const int64_t h = (l >> 1) * 4 + a.pair_slot * 2 + (l & 1);
        a.pair_w16[h] = w16;
        a.pair_ema16[h] = ne;
    }
}

// Two consecutive grid parameters (i, i + 1; i even) of one net: every input
// loaded as one vector before anything is stored, so the updates of both nets
// in one thread overlap their memory round trips.  Grid parameters only (no
// l2 term; a zero gradient leaves the parameter and its moments unchanged).
struct AdamPair {
    float2 g, w, m, v;
    uint2 st;
    h2 w16, ema;
};

// The fp16 training copy is not read: every writer of a master weight also
// writes its fp16 rounding, so the copy equals round_half(master).
__device__ __forceinline__ AdamPair adam_pair_load(const AdamArgs& a, int64_t i, bool with_grad = true) {
    AdamPair p;
    p.g = with_grad ? *(const float2*)(a.grad + i) : make_float2(0.f, 0.f);
    p.w = *(const float2*)(a.master + i);
    p.m = *(const float2*)(a.m + i);
    p.v = *(const float2*)(a.v + i);
    if (a.steps16) {
        const ushort2 st = *(const ushort2*)(a.steps16 + i);
        p.st = make_uint2(st.x, st.y);
    } else {
        p.st = *(const uint2*)(a.steps + i);
    }
    p.w16 = h2{round_half(p.w.x), round_half(p.w.y)};
    p.ema = *(const h2*)(a.ema16 + i);
    return p;
}

// One grid parameter of adam_ema_param, from loaded values.
__device__ __forceinline__ void adam_grid_update(const AdamArgs& a, float g, float& w, float& m, float& v,
                                                 uint32_t& st, _Float16& w16, _Float16& ema) {
#pragma clang fp contract(off)
    const float gr = g / a.loss_scale;
    if (gr != 0.0f) {
        const float gsq = gr * gr;
        m = a.beta1 * m + (1.0f - a.beta1) * gr;
        v = a.beta2 * v + (1.0f - a.beta2) * gsq;
        st = a.steps16 ? min(st + 1u, kShortStepMax) : st + 1u;
        float lr = a.lr;
        lr *= sqrtf(1.0f - exp2f((float)st * a.log2_beta2)) / (1.0f - exp2f((float)st * a.log2_beta1));
        const float elr = fminf(fmaxf(lr / (sqrtf(v) + a.eps), 0.0f), 3.402823466e+38f);
        w = w - elr * m;
        w16 = round_half(w);
    }
    const float e = (float)ema;
    const float wv = (float)w16;
    ema = round_half((e * a.ema_decay * a.ema_debias_old + wv * (1.0f - a.ema_decay)) * a.ema_debias_new);
}

__device__ __forceinline__ void adam_pair_step(const AdamArgs& a, AdamPair& p) {
    _Float16 w0 = p.w16[0], w1 = p.w16[1], e0 = p.ema[0], e1 = p.ema[1];
    adam_grid_update(a, p.g.x, p.w.x, p.m.x, p.v.x, p.st.x, w0, e0);
    adam_grid_update(a, p.g.y, p.w.y, p.m.y, p.v.y, p.st.y, w1, e1);
    p.w16 = h2{w0, w1};
    p.ema = h2{e0, e1};
}

__device__ __forceinline__ void adam_pair_store(const AdamArgs& a, int64_t i, const AdamPair& p) {
    *(float2*)(a.master + i) = p.w;
    *(float2*)(a.m + i) = p.m;
    *(float2*)(a.v + i) = p.v;
    if (a.steps16) *(ushort2*)(a.steps16 + i) = make_ushort2((unsigned short)p.st.x, (unsigned short)p.st.y);
    else *(uint2*)(a.steps + i) = p.st;
    *(h2*)(a.w16 + i) = p.w16;
    *(h2*)(a.ema16 + i) = p.ema;
}

__device__ void grid_adam_pair_entry(const GridAdam& ga, int64_t e, float2 gt, float2 gp) {
    const int64_t ia = ga.a.grid_begin + 2 * e, ib = ga.b.grid_begin + 2 * e;
    if (ga.keep_grad) {
        *(float2*)(const_cast<float*>(ga.a.grad) + ia) = gt;
        *(float2*)(const_cast<float*>(ga.b.grad) + ib) = gp;
    }
    AdamPair pa = adam_pair_load(ga.a, ia, false);
    AdamPair pb = adam_pair_load(ga.b, ib, false);
    pa.g = gt;
    pb.g = gp;
    adam_pair_step(ga.a, pa);
    adam_pair_step(ga.b, pb);
    adam_pair_store(ga.a, ia, pa);
    adam_pair_store(ga.b, ib, pb);
    *(h4*)(ga.a.pair_w16 + 4 * e) = h4{pa.w16[0], pa.w16[1], pb.w16[0], pb.w16[1]};
    *(h4*)(ga.a.pair_ema16 + 4 * e) = h4{pa.ema[0], pa.ema[1], pb.ema[0], pb.ema[1]};
}

// Grid parameters of the theta (a) and phi (b) nets together, one table
// entry (F = 2 parameters of each) per thread: both interleaved tables get
// whole 8-byte entries instead of one half from each net's pass.
// Blocks past the entries update the MLP parameters [0, grid_begin) of a,
// then of b (per parameter, as adam_ema_kernel).
__global__ void adam_ema_pair_kernel(const AdamArgs a, const AdamArgs b, int64_t n_entries) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t eb = (n_entries + blockDim.x - 1) / blockDim.x * blockDim.x;
    if (e >= eb) {
        int64_t i = e - eb;
        _Float16 w16, ne;
        if (i < a.grid_begin) adam_ema_param(a, i, w16, ne);
        else if ((i -= a.grid_begin) < b.grid_begin) adam_ema_param(b, i, w16, ne);
        return;
    }
    if (e >= n_entries) return;
    const int64_t ia = a.grid_begin + 2 * e, ib = b.grid_begin + 2 * e;
    AdamPair pa = adam_pair_load(a, ia);
    AdamPair pb = adam_pair_load(b, ib);
    adam_pair_step(a, pa);
    adam_pair_step(b, pb);
    adam_pair_store(a, ia, pa);
    adam_pair_store(b, ib, pb);
    *(h4*)(a.pair_w16 + 4 * e) = h4{pa.w16[0], pa.w16[1], pb.w16[0], pb.w16[1]};
    *(h4*)(a.pair_ema16 + 4 * e) = h4{pa.ema[0], pa.ema[1], pb.ema[0], pb.ema[1]};
}

void launch_adam_ema(const AdamArgs& a, hipStream_t s) {
    if (a.n_params <= 0) return;
    hipLaunchKernelGGL(adam_ema_kernel, dim3((unsigned)((a.n_params + 255) / 256)), dim3(256), 0, s, a);
}

void launch_adam_ema_pair(const AdamArgs& a, const AdamArgs& b, int64_t n_entries, hipStream_t s) {
    const int64_t blocks = (n_entries + 255) / 256 + (a.grid_begin + b.grid_begin + 255) / 256;
    if (blocks <= 0) return;
    hipLaunchKernelGGL(adam_ema_pair_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n_entries);
}

// ---------------------------------------------------------------------------
// Fragment packing.  Forward layer l holds W_l (rows = outputs); backward
// layer l holds W_l^T (rows = inputs).  K order: natural (k = 32s + 8g + j)
// when the operand comes from memory, permuted
// (k = 16(2s + j/4) + 4g + j%4) when it comes from the previous layer's
// accumulators.
// ---------------------------------------------------------------------------
// One h8 slot of a pack (which: 1 forward, 2 backward, 3 both).
__device__ __forceinline__ void pack_slot(const NetSpec& ns, const _Float16* __restrict__ w16,
                                          _Float16* __restrict__ pack, int which, int64_t slot) {
    const int L = ns.n_hidden + 1;
    const uint32_t h = (uint32_t)slot * 8;
    const uint32_t end = (which & 2) ? ns.bwd_off[L] : ns.fwd_off[L];
    const uint32_t begin = (which & 1) ? 0u : ns.bwd_off[0];
    if (h + begin >= end) return;
    const uint32_t pos = h + begin;
    bool bwd = pos >= ns.bwd_off[0];
    int l = 0;
    const uint32_t* off = bwd ? ns.bwd_off : ns.fwd_off;
    while (l + 1 < L && pos >= off[l + 1]) ++l;
    // logical matrix of this pack: rows x cols, value(m, k)
    const int H = ns.n_hidden;
    const int out_l = l == H ? ns.n_out_pad : 64;
    const int in_l = l == 0 ? ns.enc_width : 64;
    int ks, natural;
    if (!bwd) {
        ks = l == 0 ? ns.k_steps : 2;
        natural = l == 0;
    } else {
        ks = l == H ? (ns.n_out_pad + 31) / 32 : 2;
        natural = l == H;
    }
    const uint32_t local = (pos - off[l]) / 8;   // frag*64 + lane
    const int lane = local & 63;
    const int frag = local >> 6;
    const int t = frag / ks, s = frag % ks;
    const int m = 16 * t + (lane & 15);
    const int gq = lane >> 4;
    const _Float16* W = w16 + ns.layer_off[l];
    h8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = natural ? 32 * s + 8 * gq + j : 16 * (2 * s + (j >> 2)) + 4 * gq + (j & 3);
        _Float16 val = (_Float16)0.0f;
        if (!bwd) {
            if (m < out_l && k < in_l) val = W[(size_t)m * in_l + k];
        } else {
            if (m < in_l && k < out_l) val = W[(size_t)k * in_l + m];
        }
        v[j] = val;
    }
    *(h8*)(pack + pos) = v;
}

__global__ void pack_kernel(const NetSpec* __restrict__ dns, const _Float16* __restrict__ w16,
                            _Float16* __restrict__ pack, int which) {
    pack_slot(*dns, w16, pack, which, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Several packs in one launch: job blockIdx.y.
__global__ void pack_jobs_kernel(const PackJobs jobs) {
    const PackJob& j = jobs.job[blockIdx.y];
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= j.slots) return;
    pack_slot(*j.dns, j.w16, j.pack, j.which, slot);
}

int64_t pack_slots(const NetSpec& hs, int which) {
    const int L = hs.n_hidden + 1;
    const uint32_t end = (which & 2) ? hs.bwd_off[L] : hs.fwd_off[L];
    const uint32_t begin = (which & 1) ? 0u : hs.bwd_off[0];
    return (int64_t)(end - begin) / 8;
}

void launch_pack(const NetSpec* dns, const NetSpec& hs, const _Float16* w16, _Float16* pack, int which,
                 hipStream_t s) {
    const int64_t slots = pack_slots(hs, which);
    if (slots <= 0) return;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, dns, w16, pack, which);
}

void launch_pack_jobs(const PackJobs& jobs, hipStream_t s) {
    int64_t mx = 0;
    for (int k = 0; k < jobs.n; ++k) mx = jobs.job[k].slots > mx ? jobs.job[k].slots : mx;
    if (mx <= 0 || jobs.n <= 0) return;
    hipLaunchKernelGGL(pack_jobs_kernel, dim3((unsigned)((mx + 255) / 256), (unsigned)jobs.n), dim3(256), 0, s, jobs);
}

// ---------------------------------------------------------------------------
// Small utilities.
// ---------------------------------------------------------------------------
__global__ void f32_to_f16_kernel(const float* __restrict__ a, _Float16* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (_Float16)a[i];
}
__global__ void f16_to_f32_kernel(const _Float16* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void fill_u32_kernel(unsigned int* __restrict__ p, unsigned int v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_f32_kernel(float* __restrict__ p, float v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// Deterministic mean: one block, fixed-order strided partial sums then a tree.
// Mean of v[0..n) in two deterministic stages: block b sums its contiguous
// range into ws[b], then one block sums ws in order.
constexpr int kMeanBlocks = 256;

__global__ __launch_bounds__(256) void sum_part_kernel(const float* __restrict__ v, int64_t n,
                                                       float* __restrict__ ws) {
    __shared__ float red[256];
    const int64_t per = (n + kMeanBlocks - 1) / kMeanBlocks;
    const int64_t b = (int64_t)blockIdx.x * per;
    const int64_t e = b + per < n ? b + per : n;
    float s = 0.0f;
    for (int64_t i = b + threadIdx.x; i < e; i += 256) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) ws[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void mean_kernel(const float* __restrict__ ws, float scale, float* __restrict__ out) {
    __shared__ float red[256];
    red[threadIdx.x] = ws[threadIdx.x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

void launch_f32_to_f16(const float* a, _Float16* b, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(f32_to_f16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
}
void launch_f16_to_f32(const _Float16* a, float* b, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(f16_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
}
void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill_u32(unsigned int* p, unsigned int v, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}
void launch_mean(const float* v, int64_t n, float scale, float* out, float* ws, hipStream_t s) {
    hipLaunchKernelGGL(sum_part_kernel, dim3(kMeanBlocks), dim3(256), 0, s, v, n, ws);
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(256), 0, s, ws, scale, out);
}

}  // namespace dfg
