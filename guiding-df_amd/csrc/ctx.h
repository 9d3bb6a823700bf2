// Host-side state of a guide context, shared by the C-ABI translation units
// (api.cpp: networks, parameters, query/train entry points; serve.cpp:
// wavefront frames and the request server).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include <utility>

#include "../../include/dfg.h"
#include "internal.h"

namespace dfg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPX(expr)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            throw Error(e_ == hipErrorOutOfMemory ? DFG_ERR_OOM : DFG_ERR_HIP,                    \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                     \
    } while (0)

inline void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(DFG_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
inline uint32_t next_multiple(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

struct OptCfg {
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f, l2 = 1e-8f;
    float ema_decay = 0.0f;  // 0 = no EMA (inference uses the training weights)
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void ensure(size_t count) {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        HIPX(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct NetState {
    bool present = false;
    NetSpec spec;
    OptCfg opt;
    uint32_t step = 0;
    bool grad_kept = true;   // s.grad holds the whole gradient of the last step (the fused
                             // grid optimizer stores its grid part only with keep_gradients)
    DevBuf<NetSpec> d_spec;
    DevBuf<float> master, m, v, grad;
    DevBuf<_Float16> w16, ema16, pack_train, pack_infer;
    // per-parameter Adam step counters: 16-bit (saturating at 65535) when
    // both betas make the bias corrections exactly 1 from that step on,
    // else 32-bit; one of the two is allocated
    DevBuf<uint32_t> steps;
    DevBuf<uint16_t> steps16;
    bool short_steps = false;
    // training workspace
    DevBuf<_Float16> x, act, g, out, gout_rm, gout_fm;
    DevBuf<float> dx, partial, gfeat;
    // plain FP32 training workspace (feature-major fp32)
    DevBuf<float> px, pact, pout, pgout, pg;
    void release_all() {
        px.release(); pact.release(); pout.release(); pgout.release(); pg.release();
        d_spec.release(); master.release(); m.release(); v.release(); grad.release();
        w16.release(); ema16.release(); pack_train.release(); pack_infer.release(); steps.release(); steps16.release();
        x.release(); act.release(); g.release(); out.release(); gout_rm.release(); gout_fm.release();
        dx.release(); partial.release(); gfeat.release();
    }
    const _Float16* tab_infer() const { return spec.grid.present ? ema16.p + spec.grid.param_begin : nullptr; }
    const _Float16* tab_train() const { return spec.grid.present ? w16.p + spec.grid.param_begin : nullptr; }
};


}  // namespace dfg

struct dfg_ctx {
    using NetState = dfg::NetState;
    template <typename T> using DevBuf = dfg::DevBuf<T>;
    NetState net[3];
    int n_theta = 16, n_phi = 32, linear = 1, D = 9;
    uint64_t seed = 7272;
    int64_t prepass_min_rows = 32768;  // queries from which the XCD-aware grid pre-pass is used
    bool grid_bwd_bucketed = true;     // bucketed (LDS) grid backward; false: fp32 atomic scatter
    // chunked query: the paired grid pre-pass of the next row chunk runs on a
    // side stream beside the query tiles of the current one
    int query_chunks = 1;              // chunks of a large query (1: no chunking)
    int prepass_blocks_per_cu = 0;     // persistent pre-pass grid per CU in chunked mode (0: one block per item)
    int64_t pipeline_min_rows = 262144;
    int wgrad_chunk_rows = 1024;       // batch rows per weight-gradient partial sum (multiple of 64)
    bool fused_training = true;        // one-kernel forward/loss/backward/weight-gradient pass where available
    bool keep_gradients = false;       // store the paired grids' gradient although their optimizer step
                                       // runs inside the grid backward (for dfg_get_grad)
    int n_cus = 256;
    int64_t lds_per_cu = 160 * 1024;  // LDS of one CU (bytes), from the device
    int device = 0;                    // the device current at dfg_create (weights and workspace live there)
    void* noalloc_scratch = nullptr;   // that device's address of the allocation-free query's scratch array
    hipStream_t side = nullptr;
    hipEvent_t pipe_start = nullptr;
    std::vector<hipEvent_t> pipe_ev;
    // theta and phi grids of identical structure (F = 2): their inference
    // tables are also kept interleaved, one 8-byte entry per vertex for both
    bool pair_ok = false;
    DevBuf<uint2> tab_pair;        // EMA (inference) weights
    DevBuf<uint2> tab_pair_train;  // training weights (the fp16 copy of the master)
    std::string err;
    // plain-path and loss workspace
    DevBuf<float> p_enc, p_a, p_b, p_theta, loss;
    DevBuf<float> tmp, pos, mean_ws;
    // The training step's compact grid positions: the bucketed grid backward
    // sizes its slices from them in one kernel and scatters by them in
    // another, so they must not be a buffer a query also writes (queries use
    // `pos`).
    DevBuf<float> train_pos;
    DevBuf<float> train_pdf;  // fused DF training: [2][ld] |pdf| of theta and phi per row
    // training rows generated from records, and their per-step losses
    DevBuf<float> rec_feat, rec_tgt, rec_loss;
    DevBuf<uint32_t> rec_local;
    DevBuf<int64_t> rec_blk, rec_count;
    DevBuf<int2> rec_vidx;
    DevBuf<float> img_rows, img_wi2;  // rows of a pdf image
    // bucketed grid-backward workspace (shared by the nets, used in stream order)
    DevBuf<int64_t> gb_offs, gb_cursor;
    DevBuf<uint32_t> gb_idx, gb_cnt;
    DevBuf<float> gb_val;  // record values (paired backward: float4 per record)
    DevBuf<int4> gb_items;
    DevBuf<int> gb_n_items;
    DevBuf<unsigned long long> gb_part;  // fixed-point partial sums of multi-part slices
    DevBuf<int4> gb_multi;               // the multi-part slices
    // The same step scratch (loss, its mean, compact positions, the grid
    // backward's buckets) for the radiance-cache training calls, swapped in
    // for their duration (NrcScratch), so that a DF step and an NRC step
    // enqueued on two streams share no device memory.
    DevBuf<float> n_loss, n_pos, n_mean_ws, n_gb_val;
    DevBuf<int64_t> n_gb_offs, n_gb_cursor;
    DevBuf<uint32_t> n_gb_idx, n_gb_cnt;
    DevBuf<int4> n_gb_items, n_gb_multi;
    DevBuf<int> n_gb_n_items;
    DevBuf<unsigned long long> n_gb_part;
    void swap_nrc_scratch() {
        std::swap(loss, n_loss); std::swap(pos, n_pos); std::swap(mean_ws, n_mean_ws);
        std::swap(gb_val, n_gb_val); std::swap(gb_offs, n_gb_offs); std::swap(gb_cursor, n_gb_cursor);
        std::swap(gb_idx, n_gb_idx); std::swap(gb_cnt, n_gb_cnt); std::swap(gb_items, n_gb_items);
        std::swap(gb_multi, n_gb_multi); std::swap(gb_n_items, n_gb_n_items); std::swap(gb_part, n_gb_part);
    }
    ~dfg_ctx() {
        for (auto e : pipe_ev) (void)hipEventDestroy(e);
        if (pipe_start) (void)hipEventDestroy(pipe_start);
        if (side) (void)hipStreamDestroy(side);
        tab_pair.release();
        tab_pair_train.release();
        gb_offs.release(); gb_cursor.release(); gb_idx.release(); gb_cnt.release(); gb_val.release();
        gb_items.release(); gb_n_items.release(); gb_part.release(); gb_multi.release();
        for (auto& n : net) n.release_all();
        p_enc.release(); p_a.release(); p_b.release(); p_theta.release(); loss.release(); tmp.release();
        pos.release(); train_pos.release(); mean_ws.release(); train_pdf.release();
        rec_feat.release(); rec_tgt.release(); rec_loss.release(); rec_local.release(); rec_blk.release();
        rec_count.release(); rec_vidx.release(); img_rows.release(); img_wi2.release();
        n_loss.release(); n_pos.release(); n_mean_ws.release(); n_gb_val.release(); n_gb_offs.release();
        n_gb_cursor.release(); n_gb_idx.release(); n_gb_cnt.release(); n_gb_items.release(); n_gb_multi.release();
        n_gb_n_items.release(); n_gb_part.release();
    }
};

// Scope of a radiance-cache training call: its step scratch in place of the
// DF step's (host-side pointer swap; the calls are made from one host thread).
struct NrcScratch {
    dfg_ctx* c;
    explicit NrcScratch(dfg_ctx* ctx) : c(ctx) { c->swap_nrc_scratch(); }
    ~NrcScratch() { c->swap_nrc_scratch(); }
    NrcScratch(const NrcScratch&) = delete;
    NrcScratch& operator=(const NrcScratch&) = delete;
};

namespace dfg {
int fail(dfg_ctx* c, int code, const std::string& msg);
QueryArgs base_args(dfg_ctx* c);
// tiles_per_block: 64-row tiles per workgroup of the kernel the strides are for
// (4: query_fused_kernel; 1: the work-list kernel).
void fill_lds_strides(const dfg_ctx* c, QueryArgs& a, bool with_nrc, bool with_df, int tiles_per_block = 4);
void grid_prepass(dfg_ctx* c, QueryArgs& a, int64_t n, const float* feat, int stride, hipStream_t st);
// Element counts (floats) of the grid pre-pass buffers of an n-row query:
// compact positions [n][3] and each DF net's features [columns][ld].
struct PrepassSizes {
    int64_t ld = 0;
    size_t pos = 0;
    size_t gfeat[2] = {0, 0};
};
PrepassSizes prepass_sizes(const dfg_ctx* c, int64_t n);
// The grid pre-pass into caller-provided buffers of prepass_sizes(c, n).
void grid_prepass_into(dfg_ctx* c, QueryArgs& a, int64_t n, const float* feat, int stride, hipStream_t st,
                       float* pos_buf, float* gfeat_th, float* gfeat_ph);
// Resolves, once per device, the device address of the allocation-free
// query's file-scope scratch array on `device` (which must be current).
void* noalloc_scratch_setup(int device);
void check_rows(const dfg_ctx* c, int64_t n, const float* feat, int stride);
void refresh_pair_table(dfg_ctx* c, hipStream_t st);
}  // namespace dfg

#define API_BEGIN try {
#define API_END(ctx)                                                                        \
    }                                                                                       \
    catch (const dfg::Error& e) { return dfg::fail(ctx, e.code, e.what()); }                \
    catch (const std::bad_alloc&) { return dfg::fail(ctx, DFG_ERR_OOM, "host out of memory"); } \
    catch (const std::exception& e) { return dfg::fail(ctx, DFG_ERR_CONFIG, e.what()); }   \
    return DFG_OK;
