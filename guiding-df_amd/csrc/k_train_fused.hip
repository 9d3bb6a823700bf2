// Fused training pass of one MLP net: forward, loss, backward data gradients
// and weight gradients of 64-row tiles in one persistent kernel, with every
// activation and gradient of a tile kept in LDS instead of a round trip
// through HBM (the separate kernels store and re-read ~1 GB per 262k-sample
// step of the 64x5 radiance cache).  Same operations and roundings as those
// kernels (fp16 activations and gradients on MFMA with fp32 accumulation,
// the same fragment packs); only the weight gradient's summation order
// differs: each workgroup accumulates its tiles in order in MFMA registers
// and the workgroups' partial sums are added in workgroup order, so the
// result is deterministic.
// Semantics: tcnn fully_fused_mlp.cu:754-854 (forward with stored
// activations, backward, weight gradients), relative_l2_luminance.h:39-88
// (the radiance cache's loss), nrc.cpp:114-170.
//
// Workgroup: 4 waves, one 64-row tile at a time, two workgroups per CU.
// Wave w computes the forward and backward MFMAs of rows 16w..16w+15 and owns
// the weight-gradient tiles of output rows 16w..16w+15 of every layer (the
// output layer: its column tile w), accumulated over all of the workgroup's
// tiles.  LDS ([feature][row] fp16, row stride RS halves): the encoded input,
// one slot per hidden layer and the output gradient.  The backward runs layer
// by layer: the data gradient g_l of a layer is computed (masked by a_l), the
// layer's weight gradient reads g_{l+1} and a_l, and g_l then replaces a_l in
// its slot, so activations and gradients share the slots (66.8 KB for the
// 64x5 radiance cache).  The next layer's weight fragments are loaded while a
// layer computes.
#include "device.h"
#include "internal.h"

namespace dfg {

namespace {

constexpr int RS = 72;  // row stride of the LDS tiles in halves (64 rows + 8: 16-byte aligned, spreads banks)

// raw inputs of one row from the LDS feature tile ([row][16] fp32)
struct RowInL {
    const float* row;
    __device__ float operator()(int d) const { return row[d]; }
};

constexpr int kMaxIn = 16;  // raw input dims held per row (the radiance cache has 15)

// 8 consecutive K entries (features) of one row from a [feature][row] tile
__device__ __forceinline__ h8 frag_col(const _Float16* t, int k0, int row) {
    h8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = t[(k0 + j) * RS + row];
    return v;
}

// 8 consecutive rows of one feature (a 16-byte LDS read)
__device__ __forceinline__ h8 frag_row(const _Float16* t, int feature, int r0) {
    return *(const h8*)(t + feature * RS + r0);
}

// the D layout of a 64-neuron layer (lane: neurons 16T + 4g + r of row i16)
// as the next operand `act` -> [neuron][row] tile
__device__ __forceinline__ void put_act(_Float16* t, const h8 (&act)[2], int row, int g) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) t[(16 * (2 * s + (j >> 2)) + 4 * g + (j & 3)) * RS + row] = act[s][j];
}

}  // namespace

template <int H, int NT_OUT>
__global__ __launch_bounds__(256, 2) void train_fused_kernel(const TrainFusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    _Float16* sm = (_Float16*)smem_raw;
    const NetSpec& ns = *a.ns;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g = lane >> 4;
    const int KX = ns.k_steps * 32;
    const int n0t = ns.enc_width / 16;  // column tiles of the first layer's weight gradient
    constexpr int KO = (NT_OUT + 1) / 2;  // K-steps of the output gradient in the first backward product
    _Float16* X = sm;                   // [KX][RS]: encoded input a_0
    _Float16* S = X + KX * RS;          // [H][64][RS]: slot j holds a_{j+1}, then g_{j+1}
    _Float16* GO = S + H * 64 * RS;     // [16 NT_OUT][RS]: output gradient
    float* FT = (float*)(GO + 16 * NT_OUT * RS);  // [64][kMaxIn]: raw inputs of the tile's rows
    const int rw = 16 * wave + i16;     // this lane's row within the tile (forward / backward)
    const _Float16* pk = a.pack;
    const int in_dims = ns.in_dims;

    // the raw inputs and targets of a tile's rows, into registers (lanes 0..15
    // of wave w: rows 16w..16w+15), all loads in flight together, then through
    // an LDS row tile to the encoder (which indexes the inputs at run time)
    float fin[kMaxIn], tg[3];
    auto load_rows = [&](int64_t tile) {
        if (lane < 16 && tile < a.n_tiles) {
            const int64_t row = tile * 64 + rw;
            const int64_t r = row < a.n ? row : a.n - 1;
            const float* src = a.feat + r * a.stride;
#pragma unroll
            for (int d = 0; d < kMaxIn; ++d) fin[d] = d < in_dims ? src[d] : 0.0f;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) tg[ch] = a.target[r * 3 + ch];
        }
    };

    f4 acc0[8];
    f4 acch[H > 1 ? H - 1 : 1][4];
    f4 acco[NT_OUT];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc0[k] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int l = 0; l < (H > 1 ? H - 1 : 1); ++l)
#pragma unroll
        for (int k = 0; k < 4; ++k) acch[l][k] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NT_OUT; ++k) acco[k] = f4{0.f, 0.f, 0.f, 0.f};

    // one 64 x 64 layer's fragments (forward: W_l, backward: W_l^T)
    auto load_w = [&](const _Float16* base, int nt, h8 (&w)[4][2]) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < nt) {
                w[t][0] = ld_frag(base, t, 0, 2, lane);
                w[t][1] = ld_frag(base, t, 1, 2, lane);
            }
    };
    // weight gradient of layer l for the tile: dW_l += g_{l+1} a_l^T, K = the
    // tile's 64 rows; G: [out][row] tile of g_{l+1}, Ain: [in][row] tile of a_l
    auto wgrad = [&](int l, const _Float16* G, const _Float16* Ain) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int r0 = 32 * s + 8 * g;
            if (l == H) {
                const h8 ah = frag_row(Ain, 16 * wave + i16, r0);
#pragma unroll
                for (int mt = 0; mt < NT_OUT; ++mt) acco[mt] = mfma16(frag_row(G, 16 * mt + i16, r0), ah, acco[mt]);
            } else if (l == 0) {
                const h8 gr = frag_row(G, 16 * wave + i16, r0);
#pragma unroll
                for (int nt = 0; nt < 8; ++nt)
                    if (nt < n0t) acc0[nt] = mfma16(gr, frag_row(Ain, 16 * nt + i16, r0), acc0[nt]);
            } else {
                const h8 gr = frag_row(G, 16 * wave + i16, r0);
#pragma unroll
                for (int m = 1; m < H; ++m)
                    if (m == l)
#pragma unroll
                        for (int nt = 0; nt < 4; ++nt)
                            acch[m - 1][nt] = mfma16(gr, frag_row(Ain, 16 * nt + i16, r0), acch[m - 1][nt]);
            }
        }
    };

    for (int64_t tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
        const int64_t row0 = tile * 64;
        // ---- encode: lanes 0..15 of wave w encode rows 16w..16w+15 (lane = row)
        load_rows(tile);
        float tgc[3] = {tg[0], tg[1], tg[2]};
        if (lane < 16) {
            float* fr = FT + rw * kMaxIn;
#pragma unroll
            for (int d = 0; d < kMaxIn; ++d) fr[d] = fin[d];
            encode_row(ns, a.tab, RowInL{fr}, [X, rw](int c, float v) { X[c * RS + rw] = (_Float16)v; }, nullptr, 0,
                       false, true);
        }
        __syncthreads();
        // ---- forward of rows 16w..16w+15: fp16 activations (the reference's storage)
        h8 act[2];
        h8 wf[4][2];
        {
            f4 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) c[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s < ns.k_steps) {
                    const h8 b = frag_col(X, s * 32 + g * 8, rw);
#pragma unroll
                    for (int t = 0; t < 4; ++t) c[t] = mfma16(ld_frag(pk + ns.fwd_off[0], t, s, ns.k_steps, lane), b, c[t]);
                }
            load_w(pk + ns.fwd_off[1 < H ? 1 : H], 1 < H ? 4 : NT_OUT, wf);
            relu_to_act(c, act);
            put_act(S, act, rw, g);
        }
#pragma unroll
        for (int l = 1; l <= H; ++l) {
            h8 wc[4][2];
#pragma unroll
            for (int t = 0; t < 4; ++t) { wc[t][0] = wf[t][0]; wc[t][1] = wf[t][1]; }
            if (l < H) load_w(pk + ns.fwd_off[l + 1], l + 1 < H ? 4 : NT_OUT, wf);
            if (l < H) {
                f4 c[4];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    c[t] = mfma16(wc[t][1], act[1], mfma16(wc[t][0], act[0], f4{0.f, 0.f, 0.f, 0.f}));
                relu_to_act(c, act);
                put_act(S + l * 64 * RS, act, rw, g);
            } else {
                // ---- output layer and the loss: RelativeL2Luminance on the
                // fp16 prediction (outputs 0..2 sit in the lanes with g == 0:
                // outputs 4g..4g+3 of row i16)
                f4 out[NT_OUT];
#pragma unroll
                for (int t = 0; t < NT_OUT; ++t)
                    out[t] = mfma16(wc[t][1], act[1], mfma16(wc[t][0], act[0], f4{0.f, 0.f, 0.f, 0.f}));
                const int64_t row = row0 + rw;
                float gv[4] = {0.f, 0.f, 0.f, 0.f};
                if (g == 0 && row < a.n) {
#pragma clang fp contract(off)
                    const float p0 = (float)(_Float16)out[0][0], p1 = (float)(_Float16)out[0][1],
                                p2 = (float)(_Float16)out[0][2];
                    const float lum = 0.299f * p0 + 0.587f * p1 + 0.114f * p2;
                    const float den = lum * lum + 0.01f;
                    const float n_total = (float)(a.n * 3);
                    const float pc[3] = {p0, p1, p2};
                    float lsum = 0.0f;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        const float diff = pc[ch] - tgc[ch];
                        lsum += diff * diff / den / 1.0f / n_total;
                        const float grad = 2.0f * diff / den / 1.0f;
                        gv[ch] = (float)sat_h(a.loss_scale * grad / n_total);
                    }
                    a.loss[row] = lsum;
                }
                // [output][row] tile of the output gradient
#pragma unroll
                for (int t = 0; t < NT_OUT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        GO[(16 * t + 4 * g + r) * RS + rw] = t == 0 && g == 0 ? (_Float16)gv[r] : (_Float16)0.0f;
            }
        }
        // the first backward product's fragments (W_H^T, K = outputs)
        h8 wb[4][KO];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int s = 0; s < KO; ++s) wb[t][s] = ld_frag(pk + ns.bwd_off[H], t, s, KO, lane);
        __syncthreads();
        // ---- backward, layer by layer: the data gradient g_l of rows
        // 16w..16w+15 (from g_{l+1}, masked by a_l), the weight gradient of
        // layer l over all 64 rows, then g_l replaces a_l in its slot
        h8 gcur[2];
        {
            // g_H = mask(a_H) W_H^T g_out; the output gradient operand from the
            // [output][row] tile (K = outputs 8g..8g+7 of row rw)
            f4 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                c[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < KO; ++s) {
                    // outputs past 16 NT_OUT are not in the tile: zero operand
                    const int k0 = s * 32 + g * 8;
                    const h8 b = k0 < 16 * NT_OUT ? frag_col(GO, k0, rw) : h8{};
                    c[t] = mfma16(wb[t][s], b, c[t]);
                }
            }
            load_w(pk + ns.bwd_off[H - 1 >= 1 ? H - 1 : 1], 4, wf);
            const _Float16* aH = S + (H - 1) * 64 * RS;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    c[t][r] = (float)aH[(16 * t + 4 * g + r) * RS + rw] > 0.0f ? c[t][r] : 0.0f;
            c_to_act(c, gcur);
            wgrad(H, GO, aH);
            __syncthreads();
            put_act(S + (H - 1) * 64 * RS, gcur, rw, g);
            __syncthreads();
        }
#pragma unroll
        for (int l = H - 1; l >= 1; --l) {
            h8 wc[4][2];
#pragma unroll
            for (int t = 0; t < 4; ++t) { wc[t][0] = wf[t][0]; wc[t][1] = wf[t][1]; }
            if (l > 1) load_w(pk + ns.bwd_off[l - 1], 4, wf);
            f4 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                c[t] = mfma16(wc[t][1], gcur[1], mfma16(wc[t][0], gcur[0], f4{0.f, 0.f, 0.f, 0.f}));
            const _Float16* al = S + (l - 1) * 64 * RS;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    c[t][r] = (float)al[(16 * t + 4 * g + r) * RS + rw] > 0.0f ? c[t][r] : 0.0f;
            c_to_act(c, gcur);
            // dW_l = g_{l+1} a_l^T: g_{l+1} is in slot l, a_l in slot l - 1
            wgrad(l, S + l * 64 * RS, al);
            __syncthreads();
            put_act(S + (l - 1) * 64 * RS, gcur, rw, g);
            __syncthreads();
        }
        // dW_0 = g_1 x^T (g_1 in slot 0)
        wgrad(0, S, X);
        __syncthreads();
    }

    // ---- this workgroup's partial weight gradient, parameter layout [out][in]
    float* part = a.partial + (size_t)blockIdx.x * ns.n_matrix_params;
    {
        float* w0 = part + ns.layer_off[0];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt)
            if (nt < n0t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    w0[(size_t)(16 * wave + 4 * g + r) * ns.enc_width + 16 * nt + i16] = acc0[nt][r];
    }
#pragma unroll
    for (int l = 1; l < H; ++l) {
        float* wl = part + ns.layer_off[l];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) wl[(16 * wave + 4 * g + r) * 64 + 16 * nt + i16] = acch[l - 1][nt][r];
    }
    {
        float* wh = part + ns.layer_off[H];
#pragma unroll
        for (int mt = 0; mt < NT_OUT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) wh[(16 * mt + 4 * g + r) * 64 + 16 * wave + i16] = acco[mt][r];
    }
}

size_t train_fused_lds_bytes(const NetSpec& hs) {
    return (size_t)RS * 2 * (hs.k_steps * 32 + hs.n_hidden * 64 + hs.n_out_pad) + (size_t)64 * kMaxIn * 4;
}

// true when launched (a net shape this kernel is instantiated for)
bool launch_train_fused_nrc(const TrainFusedArgs& a, const NetSpec& hs, int blocks, hipStream_t s) {
    if (a.n <= 0) return true;
    if (hs.n_out_pad != 16 || hs.n_hidden != 5 || hs.k_steps > 4 || hs.enc_width % 16 || hs.in_dims > kMaxIn)
        return false;
    const size_t lds = train_fused_lds_bytes(hs);
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void*)train_fused_kernel<5, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return false;
        attr = true;
    }
    hipLaunchKernelGGL((train_fused_kernel<5, 1>), dim3((unsigned)blocks), dim3(256), lds, s, a);
    return true;
}

}  // namespace dfg
