// Allocation-free guided query: dfg_query for callers that cannot allocate
// device memory during the call (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The grid pre-pass buffers that
// dfg_query grows on demand (compact positions and both DF nets' fp32 grid
// features) come from one file-scope device array instead, whose per-device
// address is resolved once, at context creation.
//
// The array is shared by every call on a device and keeps the last call's
// contents.  Nothing in it has to start at zero: the pre-pass writes every
// position and feature of rows < n before the query kernel reads them, and
// the query kernel clamps its reads to row n - 1.  Calls that overlap on one
// device (two streams, two contexts, two host threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.h"

// 65,536 rows of the 16-level, F = 2 hash-grid pair: 12 B of positions and
// 2 x 32 x 4 B of features per row, 268 B (16.75 MiB in all; every buffer
// starts on a 256-byte boundary, which these sizes keep without padding).
static constexpr size_t kScratchBytes = size_t(65536) * (3 * 4 + 2 * 32 * 4);

// In host code `scratch` names a host-side stand-in; kernels and copies get
// the per-device address that noalloc_scratch_setup resolves.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Bytes of scratch an n-row query needs (0 below the pre-pass threshold).
size_t scratch_bytes(const dfg_ctx* c, int64_t n) {
    if (n < c->prepass_min_rows) return 0;
    const dfg::PrepassSizes z = dfg::prepass_sizes(c, n);
    return align_up(z.pos * sizeof(float)) + align_up(z.gfeat[0] * sizeof(float)) +
           align_up(z.gfeat[1] * sizeof(float));
}

// The largest multiple of 64 rows whose pre-pass fits the scratch.
int64_t max_rows(const dfg_ctx* c) {
    int64_t lo = 0, hi = (int64_t)1 << 30;
    while (hi - lo > 64) {
        const int64_t mid = (lo + hi) / 2 / 64 * 64;
        if (scratch_bytes(c, mid) <= kScratchBytes) lo = mid; else hi = mid;
    }
    return lo;
}

}  // namespace

void* dfg::noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(DFG_ERR_HIP, "device id out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(DFG_ERR_HIP, "no device address for the query scratch array");
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

extern "C" int dfg_query_noalloc(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
                                 const float* u, const uint8_t* guided, float* wi, float* nrc_rgb) {
    using namespace dfg;
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, n, feat, feat_stride);
    if (n > 0 && (!u || !wi)) throw Error(DFG_ERR_INVALID_ARG, "null u or wi");
    if (nrc_rgb && !ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (nrc_rgb && feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "NRC output needs feat_stride >= 15");
    if (!ctx->noalloc_scratch) throw Error(DFG_ERR_NOT_INIT, "query scratch not resolved for this device");
    if (scratch_bytes(ctx, n) > kScratchBytes)
        throw Error(DFG_ERR_BAD_SIZE, "n = " + std::to_string(n) + " exceeds the allocation-free maximum of " +
                                          std::to_string(max_rows(ctx)) + " rows for this configuration (" +
                                          std::to_string(kScratchBytes) + " bytes of scratch)");
    if (n == 0) return DFG_OK;
    const hipStream_t st = (hipStream_t)stream;
    QueryArgs a = base_args(ctx);
    a.feat = feat; a.stride = feat_stride; a.n = n; a.u = u; a.guided = guided; a.wi = wi; a.rgb = nrc_rgb;
    if (n >= ctx->prepass_min_rows) {
        const PrepassSizes z = prepass_sizes(ctx, n);
        unsigned char* base = (unsigned char*)ctx->noalloc_scratch;
        float* pos = (float*)base;
        float* g_th = (float*)(base + align_up(z.pos * sizeof(float)));
        float* g_ph = (float*)((unsigned char*)g_th + align_up(z.gfeat[0] * sizeof(float)));
        grid_prepass_into(ctx, a, n, feat, feat_stride, st, pos, g_th, g_ph);
    }
    fill_lds_strides(ctx, a, nrc_rgb != nullptr, true);
    launch_query(a, ctx->n_theta, ctx->n_phi, st);
    check_launch();
    API_END(ctx)
}
