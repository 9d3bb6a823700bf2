// Device side of the request server.  One round = one captured graph:
//   server_admit_kernel -> extract_pos_round_kernel -> grid pre-pass (sized on
//   the device) -> query_worklist_kernel -> server_complete_kernel.
// Admission runs here, not on the host: the first kernel of a round reads the
// request ring (host-visible pinned memory), takes requests from its head in
// arrival order while the rows in flight stay within the budget, and writes
// the round's work list.  The head position and the rows in flight are one
// 64-bit word updated by compare-and-swap, so rounds on different streams
// never take the same request and the budget holds exactly; the last kernel of
// a round gives its rows back.
#include "device.h"
#include "internal.h"

namespace dfg {

constexpr int kSuperChunk = 4;  // 256-row chunks per pre-pass item
constexpr int kCtrStride = 32;  // uint32s between the pre-pass item counters

namespace {

// Loads of host-written ring fields: atomic (never through the scalar cache),
// system scope; the tail with acquire, so the entries before it are complete.
__device__ __forceinline__ uint32_t ld_sys_u32(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int64_t ld_sys_i64(int64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ T* ld_sys_ptr(T** p) {
    return (T*)__hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

// One workgroup of 256 threads.  The CAS loop is bounded: a failed swap means
// another round (or a completion) changed the word, and the loop re-reads it;
// after kAttempts failures the round admits nothing (the host launches a
// round again later).
__global__ __launch_bounds__(256) void server_admit_kernel(ServerHostShared* hs, ServerDevState* dv, WorkList* wl,
                                                           unsigned int* counter, unsigned int* item_ctr,
                                                           AdmitParams p) {
    constexpr int kAttempts = 256;
    __shared__ int64_t n_s[DFG_MAX_SEGS];
    __shared__ int64_t tile0_s[DFG_MAX_SEGS], row0_s[DFG_MAX_SEGS];
    __shared__ uint32_t s_tail, s_head, s_k;
    __shared__ unsigned long long s_old;
    __shared__ int s_done;
    __shared__ int64_t s_rows, s_tiles;
    const int t = threadIdx.x;
    if (t == 0) {
        s_tail = ld_sys_u32(&hs->tail);
        s_done = 0;
        s_k = 0;
        s_rows = 0;
        s_tiles = 0;
    }
    __syncthreads();
    const uint32_t tail = s_tail;
    for (int attempt = 0; attempt < kAttempts; ++attempt) {
        if (t == 0) s_old = __hip_atomic_load(&dv->word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const unsigned long long old = s_old;
        const uint32_t head = (uint32_t)(old >> 32);
        uint32_t avail = tail - head;  // requests waiting (mod 2^32)
        if (avail > (uint32_t)p.max_req) avail = (uint32_t)p.max_req;
        for (uint32_t i = t; i < avail; i += blockDim.x) n_s[i] = ld_sys_i64(&hs->ring[(head + i) & (kServerRing - 1)].n);
        __syncthreads();
        if (t == 0) {
            const int64_t infl = (int64_t)(uint32_t)old;
            int64_t rows = 0, tiles = 0;
            uint32_t k = 0;
            while (k < avail) {
                const int64_t n = n_s[k];
                if (rows + n > p.round_rows) break;
                // a request larger than the budget runs alone
                if (infl + rows + n > p.budget_rows && !(k == 0 && infl == 0)) break;
                tile0_s[k] = tiles;
                row0_s[k] = rows;
                rows += n;
                tiles += (n + 63) / 64;
                ++k;
            }
            s_head = head;
            bool ok = true;
            if (k > 0) {
                const unsigned long long nw =
                    ((unsigned long long)(head + k) << 32) | (unsigned long long)(uint32_t)(infl + rows);
                ok = atomicCAS(&dv->word, old, nw) == old;
                if (ok) atomicMax(&dv->peak_rows, (unsigned long long)(infl + rows));
            }
            if (ok) {
                s_done = 1;
                s_k = k;
                s_rows = rows;
                s_tiles = tiles;
            }
        }
        __syncthreads();
        if (s_done) break;
    }
    const uint32_t k = s_done ? s_k : 0u;
    const uint32_t head = s_head;
    for (uint32_t i = t; i < k; i += blockDim.x) {
        ServerReq* r = &hs->ring[(head + i) & (kServerRing - 1)];
        WorkSeg& g = wl->seg[i];
        g.feat = ld_sys_ptr((float**)&r->feat);
        g.stride = (int)ld_sys_i64(&r->stride);
        g.u = ld_sys_ptr((float**)&r->u);
        g.guided = ld_sys_ptr((uint8_t**)&r->guided);
        g.wi = ld_sys_ptr(&r->wi);
        g.rgb = ld_sys_ptr(&r->rgb);
        g.n = n_s[i];
        g.tile_begin = tile0_s[i];
        g.gbase = row0_s[i];
    }
    if (t == 0) {
        wl->n_seg = (int)k;
        wl->total_tiles = k ? s_tiles : 0;
        wl->total_rows = k ? s_rows : 0;
        *counter = 0u;
        for (int g = 0; g < kXcdGroups; ++g) item_ctr[g * kCtrStride] = 0u;
    }
    __syncthreads();
    if (t == 0) {
        // the work list is visible to the round's later kernels by stream
        // order; the record reaches the host before its sequence number
        const uint32_t seq = dv->seq[p.slot] + 1u;
        dv->seq[p.slot] = seq;
        ServerRecord* rec = &hs->rec[p.slot];
        __hip_atomic_store(&rec->h0, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&rec->h1, head + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&rec->rows, k ? s_rows : (int64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&rec->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The round's rows leave the budget (stream-ordered after its query kernel).
__global__ void server_complete_kernel(ServerDevState* dv, const WorkList* wl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t rows = wl->total_rows;
        if (rows > 0) atomicSub(&dv->word, (unsigned long long)rows);
    }
}

// Positions of a round's rows: segment by binary search over the row offsets
// (segments are non-empty and in row order); a grid-stride loop over the row
// count the admission wrote.
__global__ void extract_pos_round_kernel(const WorkList* __restrict__ wl, int pos_dim, float* __restrict__ pos) {
    const int64_t rows = wl->total_rows;
    const int ns = wl->n_seg;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < rows; g += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wl->seg[mid].gbase <= g) lo = mid;
            else hi = mid - 1;
        }
        const WorkSeg& ws = wl->seg[lo];
        const float* f = ws.feat + (g - ws.gbase) * ws.stride + pos_dim;
        pos[3 * g + 0] = f[0];
        pos[3 * g + 1] = f[1];
        pos[3 * g + 2] = f[2];
    }
}

// The paired pre-pass with its work items counted from the device row count.
// Workgroup b runs on XCD x = b % 8 (the grid is a multiple of the XCD count)
// and serves level group x % G, like block b of the full-grid pre-pass.  The
// 8 / G XCDs of a group split its rows by 1,024-row super-chunks (XCD x takes
// every (8 / G)-th one), and the workgroups of one XCD take that XCD's items
// (phase, super-chunk) in order from the XCD's own counter: each XCD then
// holds one level at a time in its L2, as under the in-order dispatch of the
// full grid, without the drift of a static stride (which left two levels per
// L2) and without every workgroup of the chip contending for one address.
// The counters (one 128-byte line apart) start at zero: the round's admission
// clears them.
__global__ __launch_bounds__(256) void grid_encode_pair_dev_kernel(const NetSpec* __restrict__ dns, const uint2* tp,
                                                                   const float* pos, int64_t cap, int64_t ld,
                                                                   float* out_th, float* out_ph, int n_groups,
                                                                   const int64_t* n_dev, unsigned int* item_ctr) {
    __shared__ unsigned int s_item;
    const int64_t n = min(cap, *n_dev);
    if (n <= 0) return;
    const int chunks = (int)((n + 255) / 256);
    const int phases = (dns->grid.n_levels + n_groups - 1) / n_groups;
    const int x = blockIdx.x % kXcdGroups;
    const int g = x % n_groups;
    const int members = kXcdGroups / n_groups;  // XCDs per level group
    const int j = x / n_groups;                 // this XCD's share of the group's rows
    const int nsc = (chunks + kSuperChunk - 1) / kSuperChunk;
    const int my_sc = nsc > j ? (nsc - j + members - 1) / members : 0;
    const unsigned int items = (unsigned int)phases * (unsigned int)my_sc;
    unsigned int* ctr = item_ctr + x * kCtrStride;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(ctr, 1u);
        __syncthreads();
        const unsigned int it = s_item;
        __syncthreads();
        if (it >= items) break;
        const int phase = (int)(it / (unsigned int)my_sc);
        const int sc = j + members * (int)(it - (unsigned int)phase * (unsigned int)my_sc);
#pragma unroll 1
        for (int k = 0; k < kSuperChunk; ++k) {
            const int c = sc * kSuperChunk + k;
            if (c >= chunks) break;
            // the block index of the full-grid order that encodes (phase, group g, chunk c)
            grid_encode_pair_block(*dns, tp, pos, n, ld, out_th, out_ph, n_groups, chunks,
                                   phase * n_groups * chunks + c * n_groups + g);
        }
    }
}

void launch_server_admit(ServerHostShared* hs, ServerDevState* dv, WorkList* wl, unsigned int* counter,
                         unsigned int* item_ctr, const AdmitParams& p, hipStream_t s) {
    hipLaunchKernelGGL(server_admit_kernel, dim3(1), dim3(256), 0, s, hs, dv, wl, counter, item_ctr, p);
}

void launch_server_complete(ServerDevState* dv, const WorkList* wl, hipStream_t s) {
    hipLaunchKernelGGL(server_complete_kernel, dim3(1), dim3(64), 0, s, dv, wl);
}

void launch_extract_pos_round(const WorkList* wl, int pos_dim, int64_t cap_rows, float* pos, hipStream_t s) {
    int64_t blocks = (cap_rows + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(extract_pos_round_kernel, dim3((unsigned)blocks), dim3(256), 0, s, wl, pos_dim, pos);
}

void launch_grid_encode_pair_dev(const NetSpec* dns, const NetSpec& hs, const uint2* tp, const float* pos,
                                 int64_t cap, int64_t ld, float* out_th, float* out_ph, const int64_t* n_dev,
                                 unsigned int* item_ctr, int blocks, hipStream_t s) {
    if (cap <= 0 || !hs.grid.present) return;
    const int groups = hs.grid.n_levels < kPrepassLevelGroups ? hs.grid.n_levels : kPrepassLevelGroups;
    blocks = blocks / kXcdGroups * kXcdGroups;
    if (blocks < kXcdGroups) blocks = kXcdGroups;
    hipLaunchKernelGGL(grid_encode_pair_dev_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dns, tp, pos, cap, ld,
                       out_th, out_ph, groups, n_dev, item_ctr);
}

}  // namespace dfg
