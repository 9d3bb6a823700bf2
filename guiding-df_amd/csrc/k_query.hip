// Fused guided query: for a 64-row tile per wavefront,
//   theta encoding -> theta MLP (MFMA) -> theta head (softmax, inverse CDF,
//   pdf) -> phi encoding with the sampled theta -> phi MLP -> phi head ->
//   (optional) NRC encoding -> NRC MLP,
// all in one launch with the activations in registers and the encoded inputs
// and logits in LDS.  Replaces the ~10 launches and the device-wide sync per
// bounce of LearnedDistribution::inferenceStep (learnedDistribution.cpp:167-206).
#include <mutex>
#include <vector>

#include "device.h"
#include "internal.h"

namespace dfg {

struct RowIn {
    const float* row;
    int base_dims;
    float extra;
    __device__ float operator()(int d) const { return d < base_dims ? row[d] : extra; }
};

// The same inputs from registers (the row loaded at the tile's start); the
// encoders index them with a wave-uniform run-time index.
struct RowInReg {
    const float (*v)[DFG_MAX_IN];
    int base_dims;
    float extra;
    __device__ float operator()(int d) const { return d < base_dims && d < DFG_MAX_IN ? (*v)[d] : extra; }
};

// Head of one net for the lane's row: reads N logits from the LDS row, writes
// the bins back (for the dynamic reads of the sampler/pdf), optionally samples.
template <int N>
__device__ __forceinline__ float head_row(float* lrow, bool do_sample, float u, float u_extra, int linear,
                                          float x_given, float& pdf_out) {
    float z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = lrow[i];
    softmax_times_n<N>(z);
#pragma unroll
    for (int i = 0; i < N; ++i) lrow[i] = z[i];
    auto P = [&](int k) { return lrow[k]; };
    float x = x_given;
    if (do_sample) {
        if (linear) {
            const Segment sg = search_linear<N>(z, u);
            x = sample_linear<N>(sg, u, u_extra, P);
        } else {
            const Segment sg = search_nearest<N>(z, u);
            x = sample_nearest<N>(sg, u, u_extra);
        }
    }
    pdf_out = linear ? pdf_linear<N>(x, P) : pdf_nearest<N>(x, P);
    return x;
}

// One 64-row tile of a query: rows [row0, row0 + 64) of `io` (lane = row).
// LOCKSTEP: a workgroup barrier before each MLP keeps the workgroup's waves on
// the same net's weights (every wave of the workgroup must run the tile).
// PREFETCH: load each net's single grid K-step and its first layer's K-step-0
// weights ahead of its MLP.  PREFETCH_IO: load the row's raw inputs, direction
// state, uniforms and guided flag at the tile's start.  Both hold registers
// longer; the work-list kernels, nearer the register limit, take only the
// second (with the first they spill).
template <int NTH, int NPH, bool WITH_DF, bool WITH_NRC, bool LOCKSTEP = false, bool PREFETCH = false,
          bool PREFETCH_IO = PREFETCH>
__device__ __forceinline__ void query_tile(const QueryArgs& a, const RowsIO& io, int64_t row0, unsigned char* smem) {
    float* xt = (float*)smem;
    float* lo = (float*)smem + a.lo_off;
    const int lane = threadIdx.x & 63;
    const int64_t row = row0 + lane;
    const bool valid = row < io.n;
    const int64_t r = valid ? row : io.n - 1;
    const float* frow = io.feat + r * io.stride;
    float* const xrow = xt + lane * a.xs;
    auto wr = [xrow](int c, float v) { xrow[c] = v; };
    float* lrow = lo + lane * a.ls;

    W0Pre w0_nrc;  // the radiance cache's first K-step weights, loaded during phi's head
    // PREFETCH_IO: the row's raw inputs into registers at the tile's start
    // (all loads in flight together) instead of one dependent load per use
    float rin[DFG_MAX_IN];
    if constexpr (PREFETCH_IO) {
        const int nd = io.stride < DFG_MAX_IN ? io.stride : DFG_MAX_IN;
#pragma unroll
        for (int d = 0; d < DFG_MAX_IN; ++d) rin[d] = d < nd ? frow[d] : 0.0f;
    }
    auto row_in = [&](int base_dims, float extra) {
        if constexpr (PREFETCH_IO) return RowInReg{&rin, base_dims, extra};
        else return RowIn{frow, base_dims, extra};
    };
  if constexpr (WITH_DF) {
    const NetSpec& th = *a.th;
    const NetSpec& ph = *a.ph;
    const int D = th.in_dims;

    // the row's direction state, uniforms and guided flag (PREFETCH_IO:
    // loaded here, so their latency overlaps theta's encoding and MLP)
    bool guided = false;
    float4 w = make_float4(0.f, 0.f, 1.f, 1.f);
    float4 uu = make_float4(0.f, 0.f, 0.f, 0.f);
    auto load_io = [&]() {
        guided = a.wi2 == nullptr && io.u != nullptr && (io.guided == nullptr || io.guided[r] != 0);
        if (a.wi2) {
            w.x = a.wi2[2 * r];
            w.y = a.wi2[2 * r + 1];
        } else if (io.wi) {
            w = *(const float4*)(io.wi + 4 * r);
        }
        // (loaded for every row when present; only guided rows use it)
        if (a.wi2 == nullptr && io.u != nullptr) uu = *(const float4*)(io.u + 4 * r);
    };
    if constexpr (PREFETCH_IO) load_io();

    // ---- theta net (with the pre-pass, its leading grid K-steps skip the LDS tile)
    const int gk_th = a.gfeat_th ? th.grid_ksteps : 0;
    const int gk_ph = a.gfeat_ph ? ph.grid_ksteps : 0;
    // a single grid K-step (32 features) is loaded ahead: theta's before its
    // encoding, phi's before theta's head
    GridPre pre_th, pre_ph;
    const bool pth = PREFETCH && gk_th == 1, pph = PREFETCH && gk_ph == 1;
    const GridDirect gd_th{gk_th ? a.gfeat_th + io.gbase : nullptr, a.gld, gk_th, row0, io.n - 1,
                           pth ? &pre_th : nullptr};
    const GridDirect gd_ph{gk_ph ? a.gfeat_ph + io.gbase : nullptr, a.gld, gk_ph, row0, io.n - 1,
                           pph ? &pre_ph : nullptr};
    if (pth) grid_pre_load(gd_th, pre_th);
    // the first layer's K-step-0 weights of theta, loaded before its encoding
    W0Pre w0_th, w0_ph;
    if constexpr (PREFETCH) w0_pre_load(th, a.pack_th, w0_th);
    if (gk_th) {
        const int o = 32 * gk_th;
        encode_row(th, a.tab_th, row_in(D, 0.0f), [xrow, o](int c, float v) { if (c >= o) xrow[c - o] = v; },
                   nullptr, 0, true, false);
    } else {
        encode_row(th, a.tab_th, row_in(D, 0.0f), wr, a.gfeat_th ? a.gfeat_th + io.gbase + r : nullptr, a.gld,
                   false, false);
    }
    if constexpr (LOCKSTEP) __syncthreads();
    mlp_forward_tile<NTH / 16, true, 4>(th, a.pack_th, xt, a.xs, lo, a.ls, nullptr, gd_th,
                                         PREFETCH ? &w0_th : nullptr);
    if (pph) grid_pre_load(gd_ph, pre_ph);
    if constexpr (PREFETCH) w0_pre_load(ph, a.pack_ph, w0_ph);
    if (a.logits_th && valid) {
#pragma unroll
        for (int i = 0; i < NTH; ++i) a.logits_th[r * NTH + i] = lrow[i];
    }
    if constexpr (!PREFETCH_IO) load_io();
    float p_th;
    const float u0 = head_row<NTH>(lrow, guided, uu.x, uu.z, a.linear, w.x, p_th);
    const float th_in = a.theta_in ? a.theta_in[r] : theta_input(u0, NTH, a.linear);

    // ---- phi net (its own encoding, theta appended as the last raw input)
    if (gk_ph) {
        // (with a.ph_part0 > 0 the leading parts are theta's, still in the row)
        const int o = 32 * gk_ph;
        encode_row(ph, a.tab_ph, row_in(D, th_in), [xrow, o](int c, float v) { if (c >= o) xrow[c - o] = v; },
                   nullptr, 0, true, false, a.ph_part0);
    } else {
        encode_row(ph, a.tab_ph, row_in(D, th_in), wr, a.gfeat_ph ? a.gfeat_ph + io.gbase + r : nullptr, a.gld,
                   false, false);
    }
    if constexpr (LOCKSTEP) __syncthreads();
    mlp_forward_tile<NPH / 16, true, 4>(ph, a.pack_ph, xt, a.xs, lo, a.ls, nullptr, gd_ph,
                                         PREFETCH ? &w0_ph : nullptr);
    if constexpr (PREFETCH && WITH_NRC) w0_pre_load(*a.nrc, a.pack_nrc, w0_nrc);
    if (a.logits_ph && valid) {
#pragma unroll
        for (int i = 0; i < NPH; ++i) a.logits_ph[r * NPH + i] = lrow[i];
    }
    float p_ph;
    const float u1 = head_row<NPH>(lrow, guided, uu.y, uu.w, a.linear, w.y, p_ph);
    if (valid) {
        if (a.wi2) {
            // AutoRegressive::pdf, AR.h:41-43
            a.pdf_out[r] = fabsf(p_th) * fabsf(p_ph) * DFG_INV_4PI;
        } else if (io.wi) {
            *(float4*)(io.wi + 4 * r) = make_float4(u0, u1, w.z * p_th, w.w * p_ph);
        }
    }
  }

    // ---- radiance cache
    if constexpr (WITH_NRC) {
        // the radiance cache runs at the reference's precision: fp16 encoded
        // input and fp16 hidden activations (fp32 accumulation).  Its output is
        // radiance, not a pdf, and feeds training targets only.
        _Float16* const xh = (_Float16*)smem;
        _Float16* const xhrow = xh + lane * a.xs_nrc;
        encode_row(*a.nrc, a.tab_nrc, row_in(64, 0.0f), [xhrow](int c, float v) { xhrow[c] = round_half(v); },
                   nullptr, 0, false, false);
        if constexpr (LOCKSTEP) __syncthreads();
        mlp_forward_tile<1, false, 4>(*a.nrc, a.pack_nrc, xh, a.xs_nrc, lo, a.ls, nullptr,
                                      GridDirect{nullptr, 0, 0, 0, 0},
                                      (PREFETCH && WITH_DF) ? &w0_nrc : nullptr);
        if (valid && io.rgb) {
            io.rgb[3 * r + 0] = lrow[0];
            io.rgb[3 * r + 1] = lrow[1];
            io.rgb[3 * r + 2] = lrow[2];
        }
    }
}

// Bytes of LDS one 64-row tile uses (the largest of its encoded-input and
// output tiles).
__host__ __device__ inline size_t query_tile_lds(const QueryArgs& a) {
    size_t lds = (size_t)64 * a.xs * 4;
    const size_t l2 = (size_t)64 * a.ls * 4;
    const size_t l3 = (size_t)64 * a.xs_nrc * 2;
    if (l2 > lds) lds = l2;
    if (l3 > lds) lds = l3;
    return (lds + 15) / 16 * 16;
}

// Up to kQueryTilesPerBlock 64-row tiles per workgroup, one per wavefront:
// the waves of a workgroup run the same nets at about the same time (a
// barrier before each MLP), so the weight fragments one of them brings into
// the CU's L1 serve the others (each net's weights exceed the L1, so tiles in
// unrelated phases would each fetch them from L2).  The launcher picks the
// tile count that keeps the workgroup's LDS within the per-workgroup limit.
constexpr int kQueryTilesPerBlock = 4;
template <int NTH, int NPH, bool WITH_DF, bool WITH_NRC>
__global__ __launch_bounds__(64 * kQueryTilesPerBlock, 2) void query_fused_kernel(const QueryArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = threadIdx.x >> 6;
    const int tpb = blockDim.x >> 6;
    const RowsIO io{a.feat, a.stride, a.u, a.guided, a.wi, a.rgb, a.n, 0};
    query_tile<NTH, NPH, WITH_DF, WITH_NRC, true, true>(a, io, ((int64_t)blockIdx.x * tpb + w) * 64,
                                                  smem + w * query_tile_lds(a));
}

// Persistent work-list kernel: every wavefront pulls 64-row tiles from a device
// counter over a table of segments (wavefront bounces or served requests) until
// the table is exhausted.  Tiles never straddle segments, and the per-row math
// is the one of query_fused_kernel, so results do not depend on the batching.
template <int NTH, int NPH, bool WITH_NRC>
__global__ __launch_bounds__(64, 2) void query_worklist_kernel(const QueryArgs a, const WorkList* __restrict__ wl,
                                                            unsigned int* __restrict__ counter, int tiles_per_fetch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int nseg = wl->n_seg;
    const int64_t total = wl->total_tiles;
    for (;;) {
        unsigned int t0 = 0;
        if (lane == 0) t0 = atomicAdd(counter, (unsigned int)tiles_per_fetch);
        // lane 0's value, as a wave-uniform (scalar) value: the segment search
        // and the segment's pointers below then stay in scalar registers
        t0 = (unsigned int)__builtin_amdgcn_readfirstlane((int)t0);
        if ((int64_t)t0 >= total) break;
        for (int k = 0; k < tiles_per_fetch; ++k) {
            const int64_t tile = (int64_t)t0 + k;
            if (tile >= total) break;
            // the last segment starting at or before the tile (empty segments
            // share their successor's tile_begin and are skipped)
            int sg = 0, hi = nseg - 1;
            while (sg < hi) {
                const int mid = (sg + hi + 1) >> 1;
                if (wl->seg[mid].tile_begin <= tile) sg = mid;
                else hi = mid - 1;
            }
            sg = __builtin_amdgcn_readfirstlane(sg);
            const WorkSeg& ws = wl->seg[sg];
            const RowsIO io{ws.feat, ws.stride, ws.u, ws.guided, ws.wi, ws.rgb, ws.n, ws.gbase};
            query_tile<NTH, NPH, true, WITH_NRC, false, false, true>(a, io, (tile - ws.tile_begin) * 64, smem);
        }
    }
}

// Builds the work list of a wavefront frame on the device from the bounce
// sizes (so a captured graph replays with new sizes) and resets the counter.
__global__ void worklist_setup_kernel(WorkList* __restrict__ wl, const int64_t* __restrict__ sizes, int n_bounces,
                                      int64_t capacity, unsigned int* __restrict__ counter) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t tile = 0, row = 0;
    for (int b = 0; b < n_bounces; ++b) {
        int64_t nb = sizes[b];
        if (nb < 0) nb = 0;
        if (row + nb > capacity) nb = capacity - row;
        WorkSeg& s = wl->seg[b];
        s.feat = wl->seg[0].feat + row * wl->seg[0].stride;
        s.stride = wl->seg[0].stride;
        s.u = wl->seg[0].u + row * 4;
        s.guided = wl->seg[0].guided ? wl->seg[0].guided + row : nullptr;
        s.wi = wl->seg[0].wi + row * 4;
        s.rgb = wl->seg[0].rgb ? wl->seg[0].rgb + row * 3 : nullptr;
        s.n = nb;
        s.tile_begin = tile;
        s.gbase = row;
        tile += (nb + 63) / 64;
        row += nb;
    }
    wl->n_seg = n_bounces;
    wl->total_tiles = tile;
    wl->total_rows = row;
    *counter = 0u;
}



// Dynamic LDS available to one workgroup on the current device (cached per
// device).
static size_t max_block_lds() {
    static std::mutex mu;
    static std::vector<std::pair<int, size_t>> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& c : cache)
        if (c.first == dev) return c.second;
    int v = 0;
    const size_t lds =
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0 ? (size_t)v
                                                                                                          : 65536;
    cache.emplace_back(dev, lds);
    return lds;
}

// Lifts a kernel's dynamic-LDS limit to `lds` bytes when that exceeds the
// default 64 KB.  The limit is an attribute of the kernel on one device, so
// it is raised once per (kernel, device, size) (host-side cache, thread-safe).
// Returns false when the runtime refuses; the limit then stays as it was and
// a launch above it fails.
bool allow_dynamic_lds(const void* fn, size_t lds) {
    if (lds <= 65536) return true;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    static std::mutex mu;
    struct Raised { const void* fn; int dev; size_t lds; };
    static std::vector<Raised> done;
    std::lock_guard<std::mutex> lock(mu);
    Raised* have = nullptr;
    for (auto& d : done)
        if (d.fn == fn && d.dev == dev) have = &d;
    if (have && have->lds >= lds) return true;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return false;
    if (have) have->lds = lds;
    else done.push_back(Raised{fn, dev, lds});
    return true;
}

static void allow_lds(const void* fn, size_t lds) { (void)allow_dynamic_lds(fn, lds); }

template <int NTH, int NPH>
static void launch_q(const QueryArgs& a, hipStream_t s) {
    const int64_t tiles = (a.n + 63) / 64;
    const size_t tl = query_tile_lds(a);
    int tpb = kQueryTilesPerBlock;
    while (tpb > 1 && tl * tpb > max_block_lds()) tpb /= 2;
    const unsigned blocks = (unsigned)((tiles + tpb - 1) / tpb);
    const size_t lds = tl * tpb;
    const dim3 bd(64 * tpb);
    if (a.th == nullptr) {
        allow_lds((const void*)query_fused_kernel<16, 16, false, true>, lds);
        hipLaunchKernelGGL((query_fused_kernel<16, 16, false, true>), dim3(blocks), bd, lds, s, a);
    } else if (a.rgb) {
        allow_lds((const void*)query_fused_kernel<NTH, NPH, true, true>, lds);
        hipLaunchKernelGGL((query_fused_kernel<NTH, NPH, true, true>), dim3(blocks), bd, lds, s, a);
    } else {
        allow_lds((const void*)query_fused_kernel<NTH, NPH, true, false>, lds);
        hipLaunchKernelGGL((query_fused_kernel<NTH, NPH, true, false>), dim3(blocks), bd, lds, s, a);
    }
}

static size_t query_lds(const QueryArgs& a) {
    size_t lds = (size_t)64 * a.xs * 4;
    const size_t l2 = (size_t)64 * a.ls * 4;
    const size_t l3 = (size_t)64 * a.xs_nrc * 2;
    if (l2 > lds) lds = l2;
    if (l3 > lds) lds = l3;
    return lds;
}

template <int NTH, int NPH>
static void launch_wl(const QueryArgs& a, const WorkList* wl, unsigned int* counter, int blocks, int tpf,
                      hipStream_t s) {
    const size_t lds = query_lds(a);
    if (a.rgb) {
        allow_lds((const void*)query_worklist_kernel<NTH, NPH, true>, lds);
        hipLaunchKernelGGL((query_worklist_kernel<NTH, NPH, true>), dim3(blocks), dim3(64), lds, s, a, wl, counter, tpf);
    } else {
        allow_lds((const void*)query_worklist_kernel<NTH, NPH, false>, lds);
        hipLaunchKernelGGL((query_worklist_kernel<NTH, NPH, false>), dim3(blocks), dim3(64), lds, s, a, wl, counter, tpf);
    }
}

void launch_query_worklist(const QueryArgs& a, int nt, int np, const WorkList* wl, unsigned int* counter,
                           int blocks, int tiles_per_fetch, hipStream_t s) {
    if (nt == 16 && np == 32) return launch_wl<16, 32>(a, wl, counter, blocks, tiles_per_fetch, s);
    if (nt == 32 && np == 32) return launch_wl<32, 32>(a, wl, counter, blocks, tiles_per_fetch, s);
    if (nt == 64 && np == 64) return launch_wl<64, 64>(a, wl, counter, blocks, tiles_per_fetch, s);
    if (nt == 16 && np == 16) return launch_wl<16, 16>(a, wl, counter, blocks, tiles_per_fetch, s);
    if (nt == 32 && np == 64) return launch_wl<32, 64>(a, wl, counter, blocks, tiles_per_fetch, s);
}

void launch_worklist_setup(WorkList* wl, const int64_t* sizes, int n_bounces, int64_t capacity, unsigned int* counter,
                           hipStream_t s) {
    hipLaunchKernelGGL(worklist_setup_kernel, dim3(1), dim3(64), 0, s, wl, sizes, n_bounces, capacity, counter);
}

void launch_query(const QueryArgs& a, int nt, int np, hipStream_t s) {
    if (a.n <= 0) return;
    if (a.th == nullptr) return launch_q<16, 16>(a, s);  // radiance cache only
    if (nt == 16 && np == 32) return launch_q<16, 32>(a, s);
    if (nt == 32 && np == 32) return launch_q<32, 32>(a, s);
    if (nt == 64 && np == 64) return launch_q<64, 64>(a, s);
    if (nt == 16 && np == 16) return launch_q<16, 16>(a, s);
    if (nt == 32 && np == 64) return launch_q<32, 64>(a, s);
}

// ---------------------------------------------------------------------------
// XCD-aware grid encoding pre-pass (see grid_encode_block).
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(256) void grid_encode_kernel(const NetSpec* __restrict__ dns, const _Float16* table,
                                                          const float* pos, int64_t n, int64_t ld, float* out,
                                                          int n_groups, int chunks, const int64_t* n_dev) {
    grid_encode_block<F>(*dns, table, pos, n_dev ? min(n, *n_dev) : n, ld, out, n_groups, chunks);
}

// One block item (level, 256 rows) per workgroup, or, with a grid smaller than
// the item count, a persistent grid striding over the items: the grid size is
// then a multiple of the XCD count, so a workgroup keeps its XCD's level in
// every phase, and the grid holds only part of the chip's wave slots (the rest
// stays free for query tiles running beside it on another stream).
__global__ __launch_bounds__(256) void grid_encode_pair_kernel(const NetSpec* __restrict__ dns, const uint2* tp,
                                                               const float* pos, int64_t n, int64_t ld, float* out_th,
                                                               float* out_ph, int n_groups, int chunks,
                                                               const int64_t* n_dev, int items) {
    const int64_t nn = n_dev ? min(n, *n_dev) : n;
    for (int b = blockIdx.x; b < items; b += gridDim.x)
        grid_encode_pair_block(*dns, tp, pos, nn, ld, out_th, out_ph, n_groups, chunks, b);
}

// Interleaved copy {theta f0 f1, phi f0 f1} of two F = 2 grid tables.
__global__ void interleave_tables_kernel(const uint32_t* __restrict__ th, const uint32_t* __restrict__ ph,
                                         uint2* __restrict__ tp, int64_t n_entries) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n_entries) tp[e] = make_uint2(th[e], ph[e]);
}

void launch_interleave_tables(const _Float16* th, const _Float16* ph, uint2* tp, int64_t n_entries, hipStream_t s) {
    if (n_entries <= 0) return;
    hipLaunchKernelGGL(interleave_tables_kernel, dim3((unsigned)((n_entries + 255) / 256)), dim3(256), 0, s,
                       (const uint32_t*)th, (const uint32_t*)ph, tp, n_entries);
}

__global__ void extract_pos_kernel(const float* __restrict__ feat, int stride, int pos_dim, int64_t n,
                                   float* __restrict__ pos, const int64_t* n_dev) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev && *n_dev < n) n = *n_dev;
    if (i >= n * 3) return;
    const int64_t r = i / 3;
    pos[i] = feat[r * stride + pos_dim + (i - r * 3)];
}

int grid_pos_dim(const NetSpec& hs) {
    for (int i = 0; i < hs.n_parts; ++i)
        if (hs.parts[i].type == ENC_GRID) return hs.parts[i].dim_begin;
    return 0;
}

void launch_extract_pos(const float* feat, int stride, int pos_dim, int64_t n, float* pos, hipStream_t s,
                        const int64_t* n_dev) {
    if (n <= 0) return;
    hipLaunchKernelGGL(extract_pos_kernel, dim3((unsigned)((n * 3 + 255) / 256)), dim3(256), 0, s, feat, stride,
                       pos_dim, n, pos, n_dev);
}

void launch_grid_encode_pair(const NetSpec* dns, const NetSpec& hs, const uint2* tp, const float* pos, int64_t n,
                             int64_t ld, float* out_th, float* out_ph, hipStream_t s, const int64_t* n_dev,
                             int max_blocks) {
    if (n <= 0 || !hs.grid.present) return;
    const int groups = hs.grid.n_levels < kPrepassLevelGroups ? hs.grid.n_levels : kPrepassLevelGroups;
    const int phases = (hs.grid.n_levels + groups - 1) / groups;
    const int chunks = (int)((n + 255) / 256);
    const int64_t items = (int64_t)chunks * groups * phases;  // <= 2^26 for n <= 2^30
    int64_t blocks = items;
    if (max_blocks > 0 && blocks > max_blocks) blocks = max_blocks / kXcdGroups * kXcdGroups;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(grid_encode_pair_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dns, tp, pos, n, ld, out_th,
                       out_ph, groups, chunks, n_dev, (int)items);
}

void launch_grid_encode(const NetSpec* dns, const NetSpec& hs, const _Float16* table, const float* pos, int64_t n,
                        int64_t ld, float* out, hipStream_t s, const int64_t* n_dev) {
    if (n <= 0 || !hs.grid.present) return;
    const int groups = hs.grid.n_levels < kPrepassLevelGroups ? hs.grid.n_levels : kPrepassLevelGroups;
    const int phases = (hs.grid.n_levels + groups - 1) / groups;
    const int chunks = (int)((n + 255) / 256);
    const unsigned blocks = (unsigned)((int64_t)chunks * groups * phases);
    switch (hs.grid.n_feat) {
        case 1: hipLaunchKernelGGL((grid_encode_kernel<1>), dim3(blocks), dim3(256), 0, s, dns, table, pos, n, ld, out, groups, chunks, n_dev); break;
        case 2: hipLaunchKernelGGL((grid_encode_kernel<2>), dim3(blocks), dim3(256), 0, s, dns, table, pos, n, ld, out, groups, chunks, n_dev); break;
        case 4: hipLaunchKernelGGL((grid_encode_kernel<4>), dim3(blocks), dim3(256), 0, s, dns, table, pos, n, ld, out, groups, chunks, n_dev); break;
        default: hipLaunchKernelGGL((grid_encode_kernel<8>), dim3(blocks), dim3(256), 0, s, dns, table, pos, n, ld, out, groups, chunks, n_dev); break;
    }
}

// ---------------------------------------------------------------------------
// Stateless DF head on given logits (dfg_df_head).
// ---------------------------------------------------------------------------
template <int NTH, int NPH>
__global__ __launch_bounds__(64) void df_head_kernel(int64_t n, int linear, const float* __restrict__ lt,
                                                     const float* __restrict__ lp, const float* __restrict__ u,
                                                     const uint8_t* __restrict__ guided, float* __restrict__ wi,
                                                     float* __restrict__ bt, float* __restrict__ bp,
                                                     int32_t* __restrict__ seg) {
    __shared__ float srow[64][(NTH > NPH ? NTH : NPH) + 1];
    const int lane = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + lane;
    if (row >= n) return;
    float* P = srow[lane];
    const bool g = guided == nullptr || guided[row] != 0;
    float4 w = *(const float4*)(wi + 4 * row);
    const float4 uu = *(const float4*)(u + 4 * row);
    float xs[2] = {w.x, w.y};
    float pd[2];
    int sgi[2] = {-1, -1};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        constexpr int NA = NTH, NB = NPH;
        const int N = h == 0 ? NA : NB;
        const float uv = h == 0 ? uu.x : uu.y;
        const float ue = h == 0 ? uu.z : uu.w;
        auto Pf = [&](int k) { return P[k]; };
        if (h == 0) {
            float z[NA];
#pragma unroll
            for (int i = 0; i < NA; ++i) z[i] = lt[row * NA + i];
            softmax_times_n<NA>(z);
#pragma unroll
            for (int i = 0; i < NA; ++i) P[i] = z[i];
            if (bt) {
#pragma unroll
                for (int i = 0; i < NA; ++i) bt[row * NA + i] = z[i];
            }
            if (g) {
                if (linear) { Segment sg = search_linear<NA>(z, uv); sgi[0] = sg.i; xs[0] = sample_linear<NA>(sg, uv, ue, Pf); }
                else { Segment sg = search_nearest<NA>(z, uv); sgi[0] = sg.i; xs[0] = sample_nearest<NA>(sg, uv, ue); }
            }
            pd[0] = linear ? pdf_linear<NA>(xs[0], Pf) : pdf_nearest<NA>(xs[0], Pf);
        } else {
            float z[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) z[i] = lp[row * NB + i];
            softmax_times_n<NB>(z);
#pragma unroll
            for (int i = 0; i < NB; ++i) P[i] = z[i];
            if (bp) {
#pragma unroll
                for (int i = 0; i < NB; ++i) bp[row * NB + i] = z[i];
            }
            if (g) {
                if (linear) { Segment sg = search_linear<NB>(z, uv); sgi[1] = sg.i; xs[1] = sample_linear<NB>(sg, uv, ue, Pf); }
                else { Segment sg = search_nearest<NB>(z, uv); sgi[1] = sg.i; xs[1] = sample_nearest<NB>(sg, uv, ue); }
            }
            pd[1] = linear ? pdf_linear<NB>(xs[1], Pf) : pdf_nearest<NB>(xs[1], Pf);
        }
        (void)N;
    }
    *(float4*)(wi + 4 * row) = make_float4(xs[0], xs[1], w.z * pd[0], w.w * pd[1]);
    if (seg) { seg[2 * row] = sgi[0]; seg[2 * row + 1] = sgi[1]; }
}

void launch_df_head(int64_t n, int nt, int np, int linear, const float* lt, const float* lp, const float* u,
                    const uint8_t* guided, float* wi, float* bt, float* bp, int32_t* seg, hipStream_t s) {
    if (n <= 0) return;
    const unsigned blocks = (unsigned)((n + 63) / 64);
#define DFG_HEAD(A, B)                                                                                       \
    if (nt == A && np == B) {                                                                                \
        hipLaunchKernelGGL((df_head_kernel<A, B>), dim3(blocks), dim3(64), 0, s, n, linear, lt, lp, u, guided, \
                           wi, bt, bp, seg);                                                                 \
        return;                                                                                              \
    }
    DFG_HEAD(16, 16) DFG_HEAD(16, 32) DFG_HEAD(32, 32) DFG_HEAD(32, 64) DFG_HEAD(64, 64) DFG_HEAD(64, 32)
    DFG_HEAD(16, 64) DFG_HEAD(32, 16) DFG_HEAD(64, 16)
#undef DFG_HEAD
}

// ---------------------------------------------------------------------------
// generate_inference_data_AR (train.cu:334-362): shading point -> DF features.
// ---------------------------------------------------------------------------
__global__ void prepare_features_kernel(int64_t n, const float* __restrict__ pos, const float* __restrict__ aabb,
                                        const float* __restrict__ wo, const float* __restrict__ nrm,
                                        const float* __restrict__ rough, float* __restrict__ feat) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // normalizeSpatialCoord (train.h:272-276): inflate the box by 0.5% of its diagonal
    float mn[3], mx[3], d[3];
    for (int k = 0; k < 3; ++k) { mn[k] = aabb[k]; mx[k] = aabb[3 + k]; d[k] = mx[k] - mn[k]; }
    const float infl = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * 0.005f;
    float* f = feat + i * 9;
    for (int k = 0; k < 3; ++k) {
        const float lo = mn[k] - infl, hi = mx[k] + infl;
        const float c = (lo + hi) * 0.5f;
        f[k] = 0.5f + (pos[i * 3 + k] - c) / (hi - lo);
    }
    // warp_direction_for_sh (train.h:1243-1245) of the normalised wo
    const float wx = wo[i * 3], wy = wo[i * 3 + 1], wz = wo[i * 3 + 2];
    const float wl = sqrtf(wx * wx + wy * wy + wz * wz);
    f[3] = (wx / wl + 1.0f) * 0.5f;
    f[4] = (wy / wl + 1.0f) * 0.5f;
    f[5] = (wz / wl + 1.0f) * 0.5f;
    // cartesianToSphericalNormalized (math_utils.h:183-195)
    const float nx = nrm[i * 3], ny = nrm[i * 3 + 1], nz = nrm[i * 3 + 2];
    const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
    const float cz = fminf(fmaxf(nz / nl, -1.0f), 1.0f);
    float ph = atan2f(ny / nl, nx / nl);
    if (ph < 0.0f) ph += 6.28318530717958647693f;
    f[6] = acosf(cz) / 3.14159265358979323846f;
    f[7] = ph / 6.28318530717958647693f;
    // warp_roughness_for_ob (train.h:1247-1249)
    f[8] = 1.0f - expf(-rough[i]);
}

void launch_prepare_features(int64_t n, const float* pos, const float* aabb, const float* wo, const float* normal,
                             const float* rough, float* feat, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(prepare_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, pos, aabb, wo,
                       normal, rough, feat);
}

// ---------------------------------------------------------------------------
// Debug exports of one net's encoding.
// ---------------------------------------------------------------------------
__global__ void encode_debug_kernel(const NetSpec* __restrict__ dns, const _Float16* tab, const float* in, int stride,
                                    int64_t n, float* out) {
    const NetSpec& ns = *dns;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* row = in + i * stride;
    float* o = out + i * ns.enc_width;
    encode_row(ns, tab, RowIn{row, 64, 0.0f}, [&](int c, float v) {
        if (c < ns.enc_width) o[c] = v;
    });
}

void launch_encode_debug(const NetSpec* dns, int enc_width, const _Float16* tab, const float* in, int stride,
                         int64_t n, float* out, hipStream_t s) {
    (void)enc_width;
    if (n <= 0) return;
    hipLaunchKernelGGL(encode_debug_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, dns, tab, in, stride, n,
                       out);
}

__global__ void grid_index_debug_kernel(const NetSpec* __restrict__ dns, const float* in, int stride, int64_t n,
                                        uint32_t* idx) {
    const NetSpec& ns = *dns;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GridSpec& g = ns.grid;
    int pd = 0;
    for (int p = 0; p < ns.n_parts; ++p)
        if (ns.parts[p].type == ENC_GRID) pd = ns.parts[p].dim_begin;
    const float* row = in + i * stride;
    for (int l = 0; l < g.n_levels; ++l) {
        GridCorners c;
        grid_corners(g, l, row[pd], row[pd + 1], row[pd + 2], c);
        for (int k = 0; k < 8; ++k) idx[(i * g.n_levels + l) * 8 + k] = c.idx[k];
    }
}

void launch_grid_index_debug(const NetSpec* dns, const float* in, int stride, int64_t n, uint32_t* idx,
                             hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(grid_index_debug_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, dns, in, stride, n,
                       idx);
}

}  // namespace dfg
