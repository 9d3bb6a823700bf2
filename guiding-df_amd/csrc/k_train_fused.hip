// Fused training pass of one MLP net: forward, loss, backward data gradients
// and weight gradients of 64-row tiles in one persistent kernel, with every
// activation and gradient of a tile kept in LDS instead of a round trip
// through HBM (the separate kernels store and re-read ~1 GB per 262k-sample
// step of the 64x5 radiance cache, ~0.9 GB for the two 64x3 DF nets).  Same
// operations and roundings as those kernels (fp16 activations and gradients
// on MFMA with fp32 accumulation, the same fragment packs, the same softmax,
// pdf and logit-gradient arithmetic); only the weight gradient's summation
// order differs: each workgroup accumulates its tiles in order in MFMA
// registers and the workgroups' partial sums are added in workgroup order, so
// the result is deterministic.
// Semantics: tcnn fully_fused_mlp.cu:754-854 (forward with stored
// activations, backward, weight gradients), relative_l2_luminance.h:39-88
// (the radiance cache's loss), nrc.cpp:114-170; the DF nets' loss head
// combine_AR_pdf_and_dl_doutput (train.h:931-965) with the softmax Jacobian of
// AR.h:50-104 and the fp16 softmax x N of train.h:874-893.
//
// The DF loss couples the two nets through one scalar per row (the other
// net's |pdf| at the sample), so a DF step is three launches: the phi net's
// forward alone (its |pdf| per row), the theta net's training pass (reads
// phi's |pdf|, writes its own and the loss), the phi net's training pass
// (reads theta's |pdf|).
//
// Workgroup: 4 waves, one 64-row tile at a time, two workgroups per CU.
// Wave w computes the forward and backward MFMAs of rows 16w..16w+15 and owns
// the weight-gradient tiles of output rows 16w..16w+15 of every layer (the
// output layer: its column tile w), accumulated over all of the workgroup's
// tiles.  LDS ([feature][row] fp16, row stride RS halves): the encoded input,
// one slot per hidden layer and the output gradient (DF: also a [row][bin]
// fp32 tile for the softmax).  The backward runs layer by layer: the data
// gradient g_l of a layer is computed (masked by a_l), the layer's weight
// gradient reads g_{l+1} and a_l, and g_l then replaces a_l in its slot, so
// activations and gradients share the slots (66.8 KB for the 64x5 radiance
// cache, 55-60 KB for a 64x3 DF net).  The next layer's weight fragments are
// loaded while a layer computes.
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

// Loss head of the fused pass.
enum FusedLoss {
    FL_NRC = 0,       // RelativeL2Luminance on the radiance cache's RGB prediction
    FL_DF_PDF = 1,    // DF net, forward only: |pdf| of the net at the sample per row
    FL_DF_TRAIN = 2,  // DF net, full training pass with the other net's |pdf| per row
};

// N0: column tiles of 16 of the first layer's weight gradient held in registers (enc_width / 16 <= N0)
template <int H, int NT_OUT, int LOSS, int N0>
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
    constexpr int NB = 16 * NT_OUT;       // DF: bins of the net (its padded output count)
    constexpr int EW = NB + 4;            // DF: row stride of the [row][bin] fp32 tile
    // (the forward-only pass keeps no activation slots and no output gradient)
    constexpr bool FWD_ONLY = LOSS == FL_DF_PDF;
    _Float16* X = sm;                   // [KX][RS]: encoded input a_0
    _Float16* S = X + KX * RS;          // [H][64][RS]: slot j holds a_{j+1}, then g_{j+1}
    _Float16* GO = S + (FWD_ONLY ? 0 : H * 64 * RS);  // [16 NT_OUT][RS]: output gradient
    float* FT = (float*)(GO + (FWD_ONLY ? 0 : 16 * NT_OUT * RS));  // [64][kMaxIn]: raw inputs of the tile's rows
    float* BT = FT + 64 * kMaxIn;       // DF: [64][EW] softmax terms, then the bins
    const int rw = 16 * wave + i16;     // this lane's row within the tile (forward / backward)
    const _Float16* pk = a.pack;
    const int in_dims = ns.in_dims;

    // the raw inputs of a tile's rows into registers (lanes 0..15 of wave w:
    // rows 16w..16w+15), all loads in flight together, then through an LDS row
    // tile to the encoder (which indexes the inputs at run time); the targets
    // into every lane of the row (the loss head works on all 64 lanes)
    // DF: the pre-pass grid features of the tile's rows, 8 columns per lane
    // (lane (i16, g): columns 8g..8g+7 of row rw), loaded together with the
    // rows instead of one dependent load per column in the encoder
    const bool gfast = a.gfeat && a.dx_n <= 32;
    float fin[kMaxIn], tg[6], gv[8], pin = 1.0f;
    auto load_rows = [&](int64_t tile) {
        if (tile >= a.n_tiles) return;
        const int64_t row = tile * 64 + rw;
        const int64_t r = row < a.n ? row : a.n - 1;
        if constexpr (LOSS == FL_NRC) {
            if (lane < 16) {
                const float* src = a.feat + r * a.stride;
                int nd = in_dims;
                asm volatile("" : "+s"(nd));  // as below: keeps the comparisons in the tile's step
#pragma unroll
                for (int d = 0; d < kMaxIn; ++d) fin[d] = d < nd ? src[d] : 0.0f;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) tg[ch] = a.target[r * 3 + ch];
            }
        } else {
            // DF record: {theta, phi, Li rgb, wiPdf, ...} (train.cu:52-60)
            const float* t = a.target + r * 9;
#pragma unroll
            for (int ch = 0; ch < 6; ++ch) tg[ch] = t[ch];
            if (gfast)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int c = 8 * g + j;
                    gv[j] = c < a.dx_n ? a.gfeat[(size_t)c * a.ld + r] : 0.0f;
                }
            if (LOSS == FL_DF_TRAIN && row < a.n) pin = a.pdf_in[row];
            if (lane < 16) {
                const float* src = a.feat + r * a.stride;
                // the phi net's last input is theta from the record (add_theta_to_input_F)
                // (base is made opaque per tile: the compiler otherwise hoists
                // the 2 x kMaxIn comparisons below out of the tile loop and
                // keeps their masks live in scalar registers, which spill)
                int base = a.theta_in ? in_dims - 1 : in_dims;
                asm volatile("" : "+s"(base));
                const float th = a.theta_in ? theta_input(tg[0], a.theta_bins, a.linear) : 0.0f;
#pragma unroll
                for (int d = 0; d < kMaxIn; ++d) fin[d] = d < base ? src[d] : (d == base ? th : 0.0f);
            }
        }
    };

    f4 acc0[N0];
    f4 acch[H > 1 ? H - 1 : 1][4];
    f4 acco[NT_OUT];
#pragma unroll
    for (int k = 0; k < N0; ++k) acc0[k] = f4{0.f, 0.f, 0.f, 0.f};
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
                for (int nt = 0; nt < N0; ++nt)
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
        // ---- encode: lanes 0..15 of wave w put the raw inputs of rows
        // 16w..16w+15 into the row tile; each row's 4 lanes then encode its parts
        load_rows(tile);
        float* fr = FT + rw * kMaxIn;
        if (lane < 16)
#pragma unroll
            for (int d = 0; d < kMaxIn; ++d) fr[d] = fin[d];
        if constexpr (LOSS == FL_NRC) {
            // the row's nested encodings spread over its 4 lanes (lane g:
            // parts g, g + 4, ...), the raw inputs through the row tile
            __syncthreads();
            encode_row(ns, a.tab, RowInL{fr}, [X, rw](int c, float v) { X[c * RS + rw] = round_half(v); },
                       nullptr, 0, false, g == 0, g, 4);
        } else {
            // (as above; a DF row's grid features come from the pre-pass)
            __syncthreads();
            const int64_t r = row0 + rw < a.n ? row0 + rw : a.n - 1;
            encode_row(ns, a.tab, RowInL{fr}, [X, rw](int c, float v) { X[c * RS + rw] = round_half(v); },
                       a.gfeat && !gfast ? a.gfeat + r : nullptr, a.ld, gfast, g == 0, g, 4);
        }
        if (gfast)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = 8 * g + j;
                if (c < a.dx_n) X[(a.dx_begin + c) * RS + rw] = round_half(gv[j]);
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
            if constexpr (!FWD_ONLY) put_act(S, act, rw, g);
        }
        const int64_t row = row0 + rw;
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
                if constexpr (!FWD_ONLY) put_act(S + l * 64 * RS, act, rw, g);
            } else {
                f4 out[NT_OUT];
#pragma unroll
                for (int t = 0; t < NT_OUT; ++t)
                    out[t] = mfma16(wc[t][1], act[1], mfma16(wc[t][0], act[0], f4{0.f, 0.f, 0.f, 0.f}));
                if constexpr (LOSS == FL_NRC) {
                    // RelativeL2Luminance on the fp16 prediction (outputs 0..2
                    // sit in the lanes with g == 0: outputs 4g..4g+3 of row i16)
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
                            const float diff = pc[ch] - tg[ch];
                            lsum += diff * diff / den / 1.0f / n_total;
                            const float grad = 2.0f * diff / den / 1.0f;
                            gv[ch] = (float)sat_h(loss_grad(a.loss_scale * grad / n_total));
                        }
                        a.loss[row] = lsum;
                    }
                    // [output][row] tile of the output gradient
#pragma unroll
                    for (int t = 0; t < NT_OUT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            GO[(16 * t + 4 * g + r) * RS + rw] = t == 0 && g == 0 ? (_Float16)gv[r] : (_Float16)0.0f;
                } else {
#pragma clang fp contract(off)
                    // softmax x N in half (softmax_times_n_half) over the 4
                    // lanes of the row: lane (i16, g) holds bins 16t + 4g + r.
                    // The maximum by exchange (order-free), the sum of the
                    // exponentials in bin order through the row tile
                    float mx = 0.0f;
#pragma unroll
                    for (int t = 0; t < NT_OUT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r) mx = fmaxf(mx, out[t][r]);
                    mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
                    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
                    float* brow = BT + rw * EW;
#pragma unroll
                    for (int t = 0; t < NT_OUT; ++t) {
                        float4 e;
                        e.x = det_expf(out[t][0] - mx);
                        e.y = det_expf(out[t][1] - mx);
                        e.z = det_expf(out[t][2] - mx);
                        e.w = det_expf(out[t][3] - mx);
                        *(float4*)(brow + 16 * t + 4 * g) = e;
                    }
                    __syncthreads();
                    float ssum = 0.0f;
#pragma unroll
                    for (int i = 0; i < NB; i += 4) {
                        const float4 e = *(const float4*)(brow + i);
                        ssum += e.x;
                        ssum += e.y;
                        ssum += e.z;
                        ssum += e.w;
                    }
                    const float ls = det_logf(ssum);
                    float bins[NT_OUT][4];
#pragma unroll
                    for (int t = 0; t < NT_OUT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            bins[t][r] = (float)((_Float16)det_expf((out[t][r] - mx) - ls) * (_Float16)NB);
                    __syncthreads();
#pragma unroll
                    for (int t = 0; t < NT_OUT; ++t)
                        *(float4*)(brow + 16 * t + 4 * g) = make_float4(bins[t][0], bins[t][1], bins[t][2], bins[t][3]);
                    __syncthreads();
                    // the net's pdf at the sample's coordinate (sample_pdf)
                    const float x = a.theta_in ? tg[1] : tg[0];
                    auto Pb = [brow](int k) { return brow[k]; };
                    const float ap = fabsf(a.linear ? pdf_linear<NB>(x, Pb) : pdf_nearest<NB>(x, Pb));
                    if constexpr (LOSS == FL_DF_PDF) {
                        if (g == 0 && row < a.n) a.pdf_out[row] = ap;
                    } else {
                        // combine_AR_pdf_and_dl_doutput with the other net's |pdf|
                        const float other = pin;
                        const float apt = a.theta_in ? other : ap, app = a.theta_in ? ap : other;
                        const DfRowTerms rt = df_row_terms(tg, apt, app, a.loss_scale / (float)a.n);
                        if (!a.theta_in && g == 0 && row < a.n) {
                            a.loss[row] = rt.loss;
                            a.pdf_out[row] = ap;
                        }
                        const float inv_res = 1.0f / (float)NB;
                        const float multiplier = rt.prefix * other * (float)NB;
                        const DfLogitGrad<NB> dg(x, a.linear, [brow, inv_res](int k) { return brow[k] * inv_res; });
#pragma unroll
                        for (int t = 0; t < NT_OUT; ++t)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int i = 16 * t + 4 * g + r;
                                const float gi = dg(i, bins[t][r] * inv_res, multiplier);
                                GO[i * RS + rw] = row < a.n ? sat_h(loss_grad(gi)) : (_Float16)0.0f;
                            }
                    }
                }
            }
        }
        if constexpr (LOSS == FL_DF_PDF) {
            // forward only: the tiles of X and the slots are rewritten by the next tile
            __syncthreads();
            continue;
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
        // dL/dx = W_0^T g_1 for the requested input columns (the grid
        // features), fp32: lane (i16, g) holds columns 16t + 4g + r of row rw
        if (a.dx) {
            const int t_lo = a.dx_begin / 16;
            const int t_hi = (a.dx_begin + a.dx_n + 15) / 16;
            for (int t = t_lo; t < t_hi; ++t) {
                const h8 w0 = ld_frag(pk + ns.bwd_off[0], t, 0, 2, lane);
                const h8 w1 = ld_frag(pk + ns.bwd_off[0], t, 1, 2, lane);
                const f4 c = mfma16(w1, gcur[1], mfma16(w0, gcur[0], f4{0.f, 0.f, 0.f, 0.f}));
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * t + 4 * g + r - a.dx_begin;
                    if (col >= 0 && col < a.dx_n) a.dx[(size_t)col * a.ld + row] = c[r];
                }
            }
        }
        // dW_0 = g_1 x^T (g_1 in slot 0)
        wgrad(0, S, X);
        __syncthreads();
    }

    if constexpr (LOSS == FL_DF_PDF) return;
    // ---- this workgroup's partial weight gradient, parameter layout [out][in]
    float* part = a.partial + (size_t)blockIdx.x * ns.n_matrix_params;
    {
        float* w0 = part + ns.layer_off[0];
#pragma unroll
        for (int nt = 0; nt < N0; ++nt)
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

size_t train_fused_lds_bytes(const NetSpec& hs, bool df, bool fwd_only) {
    const size_t bins = df ? (size_t)64 * (hs.n_out_pad + 4) * 4 : 0;
    const size_t slots = fwd_only ? 0 : (size_t)hs.n_hidden * 64 + hs.n_out_pad;
    return (size_t)RS * 2 * (hs.k_steps * 32 + slots) + (size_t)64 * kMaxIn * 4 + bins;
}

namespace {

// one instantiation: raise its dynamic LDS limit (once per device), then launch
template <int H, int NT_OUT, int LOSS, int N0>
bool launch_fused(const TrainFusedArgs& a, size_t lds, int blocks, hipStream_t s) {
    if (!allow_dynamic_lds((const void*)train_fused_kernel<H, NT_OUT, LOSS, N0>, lds)) return false;
    hipLaunchKernelGGL((train_fused_kernel<H, NT_OUT, LOSS, N0>), dim3((unsigned)blocks), dim3(256), lds, s, a);
    return true;
}

bool fused_shape_ok(const NetSpec& hs) {
    return hs.k_steps <= 4 && hs.enc_width % 16 == 0 && hs.in_dims <= kMaxIn;
}

}  // namespace

// true when launched (a net shape this kernel is instantiated for)
bool launch_train_fused_nrc(const TrainFusedArgs& a, const NetSpec& hs, int blocks, hipStream_t s) {
    if (a.n <= 0) return true;
    if (hs.n_out_pad != 16 || hs.n_hidden != 5 || !fused_shape_ok(hs)) return false;
    return launch_fused<5, 1, FL_NRC, 8>(a, train_fused_lds_bytes(hs, false, false), blocks, s);
}

bool train_fused_df_ok(const NetSpec& hs, int n_bins) {
    return fused_shape_ok(hs) && hs.n_hidden == 3 && hs.n_out_pad == n_bins && hs.enc_width <= 96 &&
           (n_bins == 16 || n_bins == 32 || n_bins == 64);
}

bool launch_train_fused_df(const TrainFusedArgs& a, const NetSpec& hs, int n_bins, bool pdf_only, int blocks,
                           hipStream_t s) {
    if (a.n <= 0) return true;
    if (!train_fused_df_ok(hs, n_bins)) return false;
    const size_t lds = train_fused_lds_bytes(hs, true, pdf_only);
#define DFG_FUSED_DF(NT)                                                                        \
    if (n_bins == 16 * NT)                                                                      \
        return pdf_only ? launch_fused<3, NT, FL_DF_PDF, 6>(a, lds, blocks, s)                  \
                        : launch_fused<3, NT, FL_DF_TRAIN, 6>(a, lds, blocks, s);
    DFG_FUSED_DF(1)
    DFG_FUSED_DF(2)
    DFG_FUSED_DF(4)
#undef DFG_FUSED_DF
    return false;
}

}  // namespace dfg
