// Internal host-side declarations shared by the kernel translation units and
// the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "spec.h"

namespace dfg {

// Level groups of the XCD-aware grid kernels: MI355X dispatches workgroups
// round-robin over its 8 XCDs.
constexpr int kXcdGroups = 8;
// Levels in flight at once in the grid pre-pass (block b encodes level
// phase * G + b % G, so each level is served by 8 / G XCDs whose L2 holds that
// level alone).  Workgroups are dispatched in order, so an XCD on a cheap
// level cannot run ahead of one on an expensive level; with G = 8 the small
// dense levels left their XCDs idle.  Measured per 1M-query call: G = 8
// 0.851 ms, 4 0.814, 2 0.798, 1 0.812 (1: every XCD fills its L2 with every level).
constexpr int kPrepassLevelGroups = 2;

// Arguments of the fused query kernel (one launch per wavefront stage).
struct QueryArgs {
    const NetSpec* th;         // device-resident specs (see DevSpecs)
    const NetSpec* ph;
    const NetSpec* nrc;
    const _Float16* pack_th;   // inference (EMA) forward fragment packs
    const _Float16* pack_ph;
    const _Float16* pack_nrc;
    const _Float16* tab_th;    // inference (EMA) grid tables, NULL if no grid
    const _Float16* tab_ph;
    const _Float16* tab_nrc;
    const float* feat;
    int stride;
    int linear;
    int64_t n;
    const float* u;
    const uint8_t* guided;
    float* wi;
    float* rgb;          // NULL: no NRC output
    const float* wi2;    // pdf mode: given directions [n][2]; NULL otherwise
    float* pdf_out;      // pdf mode output [n]
    float* logits_th;    // debug export [n][n_out_th] or NULL
    float* logits_ph;
    const float* theta_in;  // debug: phi net theta input [n] (logits mode)
    const float* gfeat_th;   // grid features from the XCD-aware pre-pass
    const float* gfeat_ph;   // (feature-major [L*F][gld]) or NULL: gather in-kernel
    int64_t gld;
    int xs;              // LDS row strides: DF encoded input (floats),
    int xs_nrc;          // NRC encoded input (halves),
    int ls;              // logits (floats)
    int lo_off;          // logits' first float within the row (0: they alias the row's start)
    int ph_part0;        // phi's first nested part to encode: the earlier ones equal theta's,
                         // whose columns phi reads as theta left them (0: encode every part)
};

// Arguments of the training forward kernel of one net.
struct TrainFwdArgs {
    const NetSpec* ns;
    const _Float16* pack;   // training forward fragments
    const _Float16* tab;    // training grid table or NULL
    const float* feat;
    int stride;
    const float* target;    // [B][9] (phi net: theta input from dir[0]) or NULL
    int theta_bins;         // phi net: n_theta for the theta input
    int linear;
    int mode;               // 0: DF bins (fp16 softmax x N), 1: raw output (NRC)
    int64_t n;
    int64_t ld;             // padded batch
    _Float16* x;            // [k_steps*32][ld]
    _Float16* act;          // [H][64][ld]
    _Float16* out;          // row-major [ld][n_out_pad] (bins or prediction)
    const float* gfeat;     // grid features from the pre-pass or NULL
    int xs, ls;             // LDS row strides: encoded input (halves), output (floats)
};

void launch_query(const QueryArgs& a, int n_theta, int n_phi, hipStream_t s);

// Rows of one query segment (caller-owned device buffers).
struct RowsIO {
    const float* feat;
    int stride;
    const float* u;
    const uint8_t* guided;
    float* wi;
    float* rgb;
    int64_t n;
    int64_t gbase;   // row of this segment in the pre-pass feature buffers
};

// Work list of the persistent kernel: segments of rows, each starting at a
// whole tile (64 rows) of the global tile sequence.
#define DFG_MAX_SEGS 256
struct WorkSeg {
    const float* feat;
    int stride;
    const float* u;
    const uint8_t* guided;
    float* wi;
    float* rgb;
    int64_t n;
    int64_t tile_begin;
    int64_t gbase;
};
struct WorkList {
    int n_seg;
    int64_t total_tiles;
    int64_t total_rows;
    WorkSeg seg[DFG_MAX_SEGS];
};
// Raises a kernel's dynamic-LDS limit above the 64 KB default on the current
// device (once per kernel, device and size; thread-safe); false if refused.
bool allow_dynamic_lds(const void* fn, size_t lds);
void launch_query_worklist(const QueryArgs& a, int n_theta, int n_phi, const WorkList* wl, unsigned int* counter,
                           int blocks, int tiles_per_fetch, hipStream_t s);
// Frame set-up on the device: segment b = rows of bounce b packed back to back
// starting at seg[0]'s buffers; resets the tile counter.
void launch_worklist_setup(WorkList* wl, const int64_t* sizes, int n_bounces, int64_t capacity, unsigned int* counter,
                           hipStream_t s);
// Grid pre-pass on compact positions [n][3] (launch_extract_pos makes them).
// n_dev (optional, device): the live row count; rows at or beyond it are
// skipped, so a captured graph sized for the capacity n serves smaller frames.
void launch_grid_encode(const NetSpec* dns, const NetSpec& hs, const _Float16* table, const float* pos, int64_t n,
                        int64_t ld, float* out, hipStream_t s, const int64_t* n_dev = nullptr);
// Theta + phi pre-pass in one launch from their interleaved table (grids of
// identical structure, F = 2), and the interleaved copy itself.  max_blocks > 0
// caps the grid (a persistent grid of that many workgroups, rounded down to a
// multiple of the XCD count).
void launch_grid_encode_pair(const NetSpec* dns, const NetSpec& hs, const uint2* tp, const float* pos, int64_t n,
                             int64_t ld, float* out_th, float* out_ph, hipStream_t s, const int64_t* n_dev = nullptr,
                             int max_blocks = 0);
void launch_interleave_tables(const _Float16* th, const _Float16* ph, uint2* tp, int64_t n_entries, hipStream_t s);

// ---- Request server: admission on the device --------------------------------
// Queued requests live in a ring in host-visible pinned memory: the host
// fills an entry, then advances `tail` (release).  Every round of the server
// is one captured graph whose first kernel admits requests from the ring's
// head on the device (FIFO, within the row budget), builds the round's work
// list and publishes what it took in the slot's record; its last kernel
// returns the round's rows to the budget.
constexpr uint32_t kServerRing = 4096;   // ring entries (a power of two)
constexpr int kServerMaxSlots = 16;      // rounds in flight at once
struct ServerReq {                       // one queued request
    int64_t id;
    int64_t n;
    const float* feat;
    const float* u;
    const uint8_t* guided;
    float* wi;
    float* rgb;
    int64_t stride;
};
struct ServerRecord {                    // per slot: what the slot's latest round admitted
    uint32_t seq;                        // admissions on this slot so far (published last)
    uint32_t h0, h1;                     // ring positions [h0, h1) (mod 2^32)
    uint32_t pad;
    int64_t rows;
};
struct ServerHostShared {                // pinned, mapped: written by both sides
    uint32_t tail;                       // requests pushed so far (mod 2^32), host-written
    uint32_t pad[31];
    ServerRecord rec[kServerMaxSlots];   // device-written
    ServerReq ring[kServerRing];         // host-written
};
struct ServerDevState {                  // device memory
    unsigned long long word;             // (head << 32) | rows in flight (both < 2^32)
    unsigned long long peak_rows;        // largest rows in flight right after an admission
    uint32_t seq[kServerMaxSlots];
};
struct AdmitParams {
    int64_t budget_rows;                 // rows in flight at once (< 2^31)
    int64_t round_rows;                  // rows of one round (the slot's pre-pass capacity, < 2^31)
    int max_req;                         // requests of one round (<= DFG_MAX_SEGS)
    int slot;
};
// item_ctr: the round's pre-pass item counters (kXcdGroups, 32 uint32s apart), cleared here.
void launch_server_admit(ServerHostShared* hs, ServerDevState* dv, WorkList* wl, unsigned int* counter,
                         unsigned int* item_ctr, const AdmitParams& p, hipStream_t s);
void launch_server_complete(ServerDevState* dv, const WorkList* wl, hipStream_t s);
// Positions of a round's rows (device-side row count) into a compact [rows][3] array.
void launch_extract_pos_round(const WorkList* wl, int pos_dim, int64_t cap_rows, float* pos, hipStream_t s);
// Paired grid pre-pass over the first *n_dev of cap rows, with the work items
// sized on the device from *n_dev and taken in order from per-XCD counters
// (item_ctr[32 * kXcdGroups], every 32nd zero at launch) by a grid of `blocks`
// workgroups.
void launch_grid_encode_pair_dev(const NetSpec* dns, const NetSpec& hs, const uint2* tp, const float* pos,
                                 int64_t cap, int64_t ld, float* out_th, float* out_ph, const int64_t* n_dev,
                                 unsigned int* item_ctr, int blocks, hipStream_t s);
void launch_extract_pos(const float* feat, int stride, int pos_dim, int64_t n, float* pos, hipStream_t s,
                        const int64_t* n_dev = nullptr);
int grid_pos_dim(const NetSpec& hs);
// Host copies of the specs are passed beside the args where launch shapes depend on them.
void launch_train_fwd(const TrainFwdArgs& a, const NetSpec& hs, hipStream_t s);

struct TrainBwdArgs {
    const NetSpec* ns;
    const _Float16* pack;   // training packs (backward part at ns.bwd_off)
    const _Float16* gout;   // row-major [ld][n_out_pad] fp16
    const _Float16* act;    // [H][64][ld]
    _Float16* g;            // [H][64][ld]  (g_1..g_H)
    float* dx;              // [dx_n][ld] or NULL
    int dx_begin, dx_n;
    int64_t n, ld;
};
void launch_train_bwd(const TrainBwdArgs& a, const NetSpec& hs, hipStream_t s);

struct WgradArgs {
    const NetSpec* ns;
    const _Float16* x;      // [k_steps*32][ld]
    const _Float16* act;    // [H][64][ld]
    const _Float16* g;      // [H][64][ld]
    const _Float16* gout_fm;  // feature-major [n_out_pad][ld]
    float* partial;         // [chunks][n_matrix_params]
    float* grad;            // params gradient (matrix part)
    int64_t n, ld;
    int chunk;              // rows per chunk (multiple of 32)
    int n_chunks;
};
void launch_wgrad(const WgradArgs& a, const NetSpec& hs, hipStream_t s);

// Fused training pass of one net (forward, loss, backward, weight gradients
// per 64-row tile in LDS); partial weight gradients per workgroup.
struct TrainFusedArgs {
    const NetSpec* ns;
    const _Float16* pack;     // training packs (forward at fwd_off, backward at bwd_off)
    const _Float16* tab;      // training grid table or NULL
    const float* feat;
    int stride;
    const float* target;      // radiance cache: target RGB [B][3]; DF: training records [B][9]
    float* loss;              // [B] (DF: written by the theta pass)
    float* partial;           // [workgroups][n_matrix_params]
    int64_t n, n_tiles;
    float loss_scale;
    // DF nets
    const float* gfeat;       // grid features from the pre-pass, [L*F][ld] fp32, or NULL
    int64_t ld;               // padded batch (column stride of gfeat and dx)
    int theta_in;             // 1: the phi net (theta input from the record, phi coordinate)
    int theta_bins, linear;
    const float* pdf_in;      // [B] the other net's |pdf| at the sample (training pass)
    float* pdf_out;           // [B] this net's |pdf| (pdf pass; theta's training pass)
    float* dx;                // [dx_n][ld] dL/d(grid features) or NULL
    int dx_begin, dx_n;       // the grid features' columns in the encoded input (and dx's columns)
};
size_t train_fused_lds_bytes(const NetSpec& hs, bool df, bool fwd_only);
bool launch_train_fused_nrc(const TrainFusedArgs& a, const NetSpec& hs, int blocks, hipStream_t s);
// DF nets of 3 hidden layers and 16/32/64 bins: the forward-only |pdf| pass
// (pdf_only) or the training pass; false when the shape is not instantiated.
bool train_fused_df_ok(const NetSpec& hs, int n_bins);
bool launch_train_fused_df(const TrainFusedArgs& a, const NetSpec& hs, int n_bins, bool pdf_only, int blocks,
                           hipStream_t s);
void launch_wgrad_reduce(const float* partial, int n_chunks, int64_t n_matrix, float* grad, hipStream_t s);

struct DfLossArgs {
    const float* target;    // [B][9]
    const _Float16* bins_th;  // [ld][nt_pad]
    const _Float16* bins_ph;
    int nt_pad, np_pad;
    _Float16* gout_th_rm;   // row-major [ld][nt_pad]
    _Float16* gout_ph_rm;
    _Float16* gout_th_fm;   // feature-major [nt_pad][ld]
    _Float16* gout_ph_fm;
    float* loss;            // [B]
    int64_t n, ld;
    float loss_scale;
    int linear;
};
void launch_df_loss(const DfLossArgs& a, int n_theta, int n_phi, hipStream_t s);

struct NrcLossArgs {
    const _Float16* pred;   // [ld][16]
    const float* target;    // [B][3]
    _Float16* gout_rm;      // [ld][16]
    _Float16* gout_fm;      // [16][ld]
    float* loss;            // [B]
    int64_t n, ld;
    float loss_scale;
};
void launch_nrc_loss(const NrcLossArgs& a, hipStream_t s);

struct GridBwdArgs {
    const NetSpec* ns;
    const float* feat;
    int stride;
    int pos_dim;            // first raw input dim of the grid
    const float* dx;        // [L*F][ld] dL/d(grid features)
    float* grad;            // parameter gradient base (grid part at ns.grid.param_begin)
    int64_t n, ld;
};
void launch_grid_bwd(const GridBwdArgs& a, const NetSpec& hs, hipStream_t s);

// Bucketed grid backward: the table gradient is built slice by slice in LDS
// (one slice = kGridSliceFloats gradient values of one level) from compact
// update records, as 64-bit fixed-point sums, then written once; no global
// atomics, no zero fill.
constexpr int kGridSliceFloats = 8192;        // 64 KB of int64 LDS accumulators per slice
constexpr int kGridMaxSlicesPerLevel = 2048;
constexpr int kGridChunkRows = 4096;
constexpr int kPairGroupBytes = 144;          // paired records: 8 float4 values, then their 8 16-bit entries
constexpr int kGridPartRecords = 65536;       // records reduced by one block
struct GridBucketPlan {
    bool ok;                 // the configuration fits the bucketed kernels
    int slice_shift;         // log2 entries per slice
    int n_slices;            // all levels
    int n_chunks;            // row chunks of kGridChunkRows
    int max_level_slices;    // largest slice count of one level
    int64_t records;         // n * levels * corners
    int64_t max_items;       // bound on (slice, part) work items
    int64_t max_partials;    // bound on the parts of slices cut into several parts
    int64_t max_multi;       // bound on the slices cut into several parts
};
// values_per_entry: the grid's F, or 4 for the paired theta/phi records.
GridBucketPlan grid_bucket_plan(const NetSpec& hs, int64_t n, int values_per_entry = 0);
struct GridBucketArgs {
    const NetSpec* ns;
    const float* feat;
    int stride;
    int pos_dim;
    const float* dx;         // [L*F][ld]
    float* grad;             // parameter gradient base
    int64_t n, ld;
    int slice_shift;
    int n_chunks;
    uint32_t* cnt;           // [n_slices][n_chunks] updates per slice and row chunk
    int64_t* offs;           // [n_slices + 1]: slice totals, then exclusive offsets
    int64_t* cursor;         // [n_slices] next free record of each slice
    uint32_t* rec_idx;       // [records] entry within the slice (F != 2, unpaired)
    float* rec_val;          // F == 2: [records] 12-byte {entry, g0, g1}; paired: groups of 8
                             // records (kPairGroupBytes); else [records][F]
    int4* items;             // [max_items] (slice, part, parts of the slice, partial slot)
    int* n_items;            // [2]: work items, slices cut into several parts
    int4* multi;             // [max_multi] (slice, first partial slot, parts, -)
    uint32_t* lvl_max;       // [levels] bits of max |update| per level (zeroed with offs)
    unsigned long long* partial;  // [max_partials][kGridSliceFloats] fixed-point sums of
                                  // the parts of slices cut into several parts
    // second net of the paired backward (same grid structure): its dL/dy,
    // gradient, first grid parameter and level maxima; the records are then
    // 18 bytes, in groups of 8 in rec_val: the 8 float4 {g0, g1, g0', g1'},
    // then the 8 16-bit entries in the slice (one contiguous 144-byte run)
    const float* dx2;
    float* grad2;
    int64_t param_begin2;
    uint32_t* lvl_max2;
};
void launch_grid_bwd_bucketed(const GridBucketArgs& a, const NetSpec& hs, const GridBucketPlan& p, hipStream_t s);
// Theta and phi together (F = 2 grids of identical structure, not nearest):
// a.*2 describe phi; p from grid_bucket_plan(hs, n, 4).


struct AdamArgs {
    int64_t n_params, n_matrix;
    float lr, beta1, beta2, eps, l2_reg, loss_scale;
    float log2_beta1, log2_beta2;
    // optional interleaved theta/phi tables (F = 2 grids of identical
    // structure): net `pair_slot` writes its half of each entry
    _Float16* pair_w16;
    _Float16* pair_ema16;
    int64_t grid_begin;
    int pair_slot;
    float ema_decay, ema_debias_old, ema_debias_new;
    float* master;
    _Float16* w16;
    _Float16* ema16;
    const float* grad;
    float* m;
    float* v;
    uint32_t* steps;     // per-parameter step counters, or NULL when steps16 is used
    uint16_t* steps16;   // 16-bit counters saturating at kShortStepMax
};
// A 16-bit counter saturates here.  It is used only when beta^t rounds 1 - beta^t
// to 1.0f for both betas at t = kShortStepMax (and so for every later t), so
// the saturated counter gives the same bias corrections as the exact one.
constexpr uint32_t kShortStepMax = 65535u;
void launch_adam_ema(const AdamArgs& a, hipStream_t s);
// Grid parameters [grid_begin, grid_begin + 2 n_entries) of two nets whose
// grids share the interleaved tables (a: theta, b: phi), whole entries per
// thread, and in the same launch their MLP parameters [0, grid_begin).
void launch_adam_ema_pair(const AdamArgs& a, const AdamArgs& b, int64_t n_entries, hipStream_t s);

// Adam + EMA of the paired grids' parameters fused into the paired grid
// backward: each slice's fixed-point sums update the parameters of its
// entries directly (both nets, whole interleaved-table entries); the gradient
// is stored only with keep_grad.  on == 0: the reduce stores the gradient and
// the optimizer runs separately.
struct GridAdam {
    AdamArgs a, b;
    int on, keep_grad;
};
void launch_grid_bwd_bucketed_pair(const GridBucketArgs& a, const NetSpec& hs, const GridBucketPlan& p,
                                   const GridAdam& ga, hipStream_t s);

// Pack fragments from fp16 parameters. which: 1 forward, 2 backward, 3 both.
void launch_pack(const NetSpec* dns, const NetSpec& hs, const _Float16* w16, _Float16* pack, int which, hipStream_t s);
int64_t pack_slots(const NetSpec& hs, int which);
// Up to four packs in one launch.
// sampleWi (integrator.cpp:355-401): the caller's uniforms, BSDF sample and
// smooth flags in; the RR / guide / mixture flags and wi slots out.
struct SampleWiArgs {
    float rr_prob, bsdf_fraction;
    int guiding_enabled, max_guided_depth;
    const int32_t* depth;
    const uint8_t* smooth;
    const float* u_rr;
    const float* u_mix;
    const float* bsdf_wi;
    const float* bsdf_pdf;
    const float* bsdf_f;
    uint8_t* rr_killed;
    uint8_t* guided;
    uint8_t* mix;
    float* wi;
    float* thp;
};

// The training-record step of handleIntersections (integrator.cpp:277-296).
struct RecordScatterArgs {
    int train_enabled, max_train_depth, mis_aware, product_with_cosine;
    uint32_t pixel_offset, pixel_stride;
    const int32_t* pixel;
    const int32_t* depth;
    const uint8_t* rr_killed;
    const uint8_t* alive;
    const float* wi_dir;
    const float* thp;
    const float* scatter_pdf;
    const float* mis;
    const float* bsdf_val;
    const float* cos_theta;
    const uint8_t* delta;
};

// recordRadiance / overwriteLe of handleEmissiveHit and handleMiss
// (integrator.cpp:144-191).
struct RecordRadianceArgs {
    int train_enabled, max_train_depth;
    uint32_t pixel_offset, pixel_stride;
    const int32_t* pixel;
    const int32_t* depth;
    const float* contrib;
    const float* Le;
    const uint8_t* miss;
};

struct PackJob {
    const NetSpec* dns;
    const _Float16* w16;
    _Float16* pack;
    int which;
    int64_t slots;  // pack_slots(host spec, which)
};
struct PackJobs {
    PackJob job[4];
    int n;
};
void launch_pack_jobs(const PackJobs& jobs, hipStream_t s);
void launch_f32_to_f16(const float* src, _Float16* dst, int64_t n, hipStream_t s);
void launch_f16_to_f32(const _Float16* src, float* dst, int64_t n, hipStream_t s);
void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s);
void launch_fill_u32(unsigned int* p, unsigned int v, int64_t n, hipStream_t s);
// Mean of v[0..n) times scale into out[0]; ws holds 256 floats of partial sums.
void launch_mean(const float* v, int64_t n, float scale, float* out, float* ws, hipStream_t s);

// Plain FP32 per-batch path.
void launch_plain_encode(const NetSpec* dns, const _Float16* tab, const float* feat, int stride,
                         const float* extra, int64_t n, int64_t ld, float* out, hipStream_t s);
void launch_plain_layer(const _Float16* W, int M, int K, int K_alloc, const float* X, float* Y,
                        int relu, int64_t n, int64_t ld, hipStream_t s);
void launch_plain_head(int phase, int n_bins, int n_theta, int linear, const float* logits,
                       int64_t ld, int64_t n, const float* u, const uint8_t* guided, float* wi,
                       float* theta_in, hipStream_t s);
void launch_plain_rgb(const float* out, int64_t ld, int64_t n, float* rgb, hipStream_t s);
// Plain FP32 training step (the comparator of the training path).
void launch_wgrad_reduce(const float* partial, int n_chunks, int64_t n_matrix, float* grad, hipStream_t s);
void launch_plain_theta_target(const float* target, int64_t n, int n_theta, int linear, float* out, hipStream_t s);
void launch_plain_layer_bwd(const _Float16* W, int K_row, int m_begin, int M, int K, const float* dY,
                            const float* mask, float* dX, int64_t n, int64_t ld, hipStream_t s);
void launch_plain_df_loss(const float* target, const float* lt, const float* lp, int n_theta, int n_phi,
                          int linear, float* g_th, float* g_ph, float* loss, float loss_scale, int64_t n,
                          int64_t ld, hipStream_t s);
void launch_plain_nrc_loss(const float* pred, const float* target, float* g, float* loss, float loss_scale,
                           int64_t n, int64_t ld, hipStream_t s);
void launch_plain_wgrad(const float* dY, int M, const float* X, int K, float* partial, int64_t pstride,
                        int64_t n, int64_t ld, int chunk, int n_chunks, hipStream_t s);

// Stateless DF head.
void launch_df_head(int64_t n, int n_theta, int n_phi, int linear, const float* lt,
                    const float* lp, const float* u, const uint8_t* guided, float* wi,
                    float* bt, float* bp, int32_t* seg, hipStream_t s);
void launch_prepare_features(int64_t n, const float* pos, const float* aabb, const float* wo,
                             const float* normal, const float* rough, float* feat, hipStream_t s);
// Debug encode / grid indices for one net.
void launch_encode_debug(const NetSpec* dns, int enc_width, const _Float16* tab, const float* in, int stride,
                         int64_t n, float* out, hipStream_t s);
void launch_grid_index_debug(const NetSpec* dns, const float* in, int stride, int64_t n,
                             uint32_t* idx, hipStream_t s);

}  // namespace dfg
