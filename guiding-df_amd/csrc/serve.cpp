// Batched wavefront frames (persistent work-list kernel in a captured
// hipGraph, autotuned) and the continuous-batching request server.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstddef>
#include <deque>
#include <memory>
#include <vector>

#include "ctx.h"

using namespace dfg;

namespace {

int cu_count() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}

}  // namespace

// ============================================================================
// Wavefront frames
// ============================================================================
struct dfg_wavefront {
    dfg_ctx* ctx = nullptr;
    int max_bounces = 0;
    int64_t capacity = 0;
    dfg_wavefront_tuning tune{4, 1, 1, 0.0f};
    DevBuf<WorkList> wl;
    DevBuf<unsigned int> counter;
    DevBuf<float> pos;
    DevBuf<float> gth, gph;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // what the graph was captured for
    const int64_t* c_sizes = nullptr;
    const float* c_feat = nullptr;
    int c_stride = 0;
    const float* c_u = nullptr;
    const uint8_t* c_guided = nullptr;
    float* c_wi = nullptr;
    int c_bounces = -1;
    dfg_wavefront_tuning c_tune{0, 0, -1, 0.0f};

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~dfg_wavefront() {
        drop_graph();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        wl.release(); counter.release(); pos.release(); gth.release(); gph.release();
    }
};

static void wavefront_enqueue(dfg_wavefront* w, hipStream_t st, int n_bounces, const int64_t* sizes) {
    dfg_ctx* c = w->ctx;
    const int64_t ld = round_up(w->capacity, 64);
    launch_worklist_setup(w->wl.p, sizes, n_bounces, w->capacity, w->counter.p, st);
    QueryArgs a = base_args(c);
    a.feat = w->c_feat;
    a.stride = w->c_stride;
    a.n = w->capacity;
    const int64_t* n_dev = &w->wl.p->total_rows;
    if (w->tune.grid_prepass && c->pair_ok) {
        launch_extract_pos(w->c_feat, w->c_stride, grid_pos_dim(c->net[0].spec), w->capacity, w->pos.p, st, n_dev);
        launch_grid_encode_pair(c->net[0].d_spec.p, c->net[0].spec, c->tab_pair.p, w->pos.p, w->capacity, ld,
                                w->gth.p, w->gph.p, st, n_dev);
        a.gfeat_th = w->gth.p;
        a.gfeat_ph = w->gph.p;
        a.gld = ld;
    } else if (w->tune.grid_prepass) {
        int pd = -1;
        for (int i = 0; i < 2; ++i) {
            NetState& s = c->net[i];
            if (!s.spec.grid.present) continue;
            const int d = grid_pos_dim(s.spec);
            if (d != pd) {
                launch_extract_pos(w->c_feat, w->c_stride, d, w->capacity, w->pos.p, st, n_dev);
                pd = d;
            }
            float* out = i == 0 ? w->gth.p : w->gph.p;
            launch_grid_encode(s.d_spec.p, s.spec, s.tab_infer(), w->pos.p, w->capacity, ld, out, st, n_dev);
            (i == 0 ? a.gfeat_th : a.gfeat_ph) = out;
        }
        a.gld = ld;
    }
    fill_lds_strides(c, a, false, true);
    const int blocks = std::max(1, cu_count() * std::max(1, w->tune.blocks_per_cu));
    launch_query_worklist(a, c->n_theta, c->n_phi, w->wl.p, w->counter.p, blocks, std::max(1, w->tune.tiles_per_fetch),
                          st);
}

static void wavefront_capture(dfg_wavefront* w, int n_bounces, const int64_t* sizes) {
    w->drop_graph();
    // base pointers of the packed frame (segment 0) for the device-side set-up
    WorkList head;
    std::memset(&head, 0, sizeof(head));
    WorkSeg& s0 = head.seg[0];
    s0.feat = w->c_feat;
    s0.stride = w->c_stride;
    s0.u = w->c_u;
    s0.guided = w->c_guided;
    s0.wi = w->c_wi;
    HIPX(hipMemcpy(w->wl.p, &head, offsetof(WorkList, seg) + sizeof(WorkSeg), hipMemcpyHostToDevice));
    HIPX(hipStreamBeginCapture(w->cap_stream, hipStreamCaptureModeThreadLocal));
    wavefront_enqueue(w, w->cap_stream, n_bounces, sizes);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(w->cap_stream, &g);
    if (e != hipSuccess) throw Error(DFG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    w->graph = g;
    HIPX(hipGraphInstantiate(&w->exec, g, nullptr, nullptr, 0));
    w->c_sizes = sizes;
    w->c_bounces = n_bounces;
    w->c_tune = w->tune;
}

static void wavefront_check(dfg_wavefront* w, int n_bounces, const int64_t* sizes, const float* feat, int stride,
                            const float* u, float* wi) {
    if (n_bounces < 1 || n_bounces > w->max_bounces) throw Error(DFG_ERR_BAD_SIZE, "n_bounces out of range");
    if (!sizes || !feat || !u || !wi) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (stride < w->ctx->D) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below n_input_dims");
}

static void wavefront_run(dfg_wavefront* w, hipStream_t st, int n_bounces, const int64_t* sizes, const float* feat,
                          int stride, const float* u, const uint8_t* guided, float* wi) {
    const bool same = w->exec && w->c_sizes == sizes && w->c_feat == feat && w->c_stride == stride && w->c_u == u &&
                      w->c_guided == guided && w->c_wi == wi && w->c_bounces == n_bounces &&
                      w->c_tune.blocks_per_cu == w->tune.blocks_per_cu &&
                      w->c_tune.tiles_per_fetch == w->tune.tiles_per_fetch &&
                      w->c_tune.grid_prepass == w->tune.grid_prepass;
    if (!same) {
        w->c_feat = feat;
        w->c_stride = stride;
        w->c_u = u;
        w->c_guided = guided;
        w->c_wi = wi;
        wavefront_capture(w, n_bounces, sizes);
    }
    HIPX(hipGraphLaunch(w->exec, st));
}

extern "C" {

int dfg_wavefront_create(dfg_ctx* ctx, int max_bounces, int64_t capacity, dfg_wavefront** out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!out) throw Error(DFG_ERR_INVALID_ARG, "null output");
    if (max_bounces < 1 || max_bounces > DFG_MAX_SEGS) throw Error(DFG_ERR_BAD_SIZE, "max_bounces must be 1..256");
    if (capacity < 1 || capacity > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "capacity out of range");
    std::unique_ptr<dfg_wavefront> w(new dfg_wavefront());
    w->ctx = ctx;
    w->max_bounces = max_bounces;
    w->capacity = capacity;
    w->wl.ensure(1);
    w->counter.ensure(1);
    const int64_t ld = round_up(capacity, 64);
    w->pos.ensure((size_t)capacity * 3);
    for (int i = 0; i < 2; ++i) {
        const NetSpec& s = ctx->net[i].spec;
        if (!s.grid.present) continue;
        (i == 0 ? w->gth : w->gph).ensure((size_t)s.grid.n_levels * s.grid.n_feat * ld);
    }
    HIPX(hipStreamCreateWithFlags(&w->cap_stream, hipStreamNonBlocking));
    *out = w.release();
    API_END(ctx)
}

int dfg_wavefront_destroy(dfg_wavefront* w) {
    if (!w) return DFG_ERR_NOT_INIT;
    delete w;
    return DFG_OK;
}

int dfg_wavefront_set_tuning(dfg_wavefront* w, const dfg_wavefront_tuning* t) {
    if (!w) return DFG_ERR_NOT_INIT;
    if (!t || t->blocks_per_cu < 1 || t->blocks_per_cu > 64 || t->tiles_per_fetch < 1 || t->tiles_per_fetch > 64)
        return fail(w->ctx, DFG_ERR_INVALID_ARG, "bad tuning");
    w->tune = *t;
    return DFG_OK;
}

int dfg_wavefront_get_tuning(const dfg_wavefront* w, dfg_wavefront_tuning* t) {
    if (!w) return DFG_ERR_NOT_INIT;
    if (!t) return DFG_ERR_INVALID_ARG;
    *t = w->tune;
    return DFG_OK;
}

int dfg_wavefront_run(dfg_wavefront* w, void* stream, int n_bounces, const int64_t* sizes, const float* feat,
                      int feat_stride, const float* u, const uint8_t* guided, float* wi) {
    if (!w) return DFG_ERR_NOT_INIT;
    dfg_ctx* ctx = w->ctx;
    API_BEGIN
    wavefront_check(w, n_bounces, sizes, feat, feat_stride, u, wi);
    wavefront_run(w, (hipStream_t)stream, n_bounces, sizes, feat, feat_stride, u, guided, wi);
    API_END(ctx)
}

int dfg_wavefront_autotune(dfg_wavefront* w, void* stream, int n_bounces, const int64_t* sizes, const float* feat,
                           int feat_stride, const float* u, const uint8_t* guided, float* wi,
                           dfg_wavefront_tuning* best) {
    if (!w) return DFG_ERR_NOT_INIT;
    dfg_ctx* ctx = w->ctx;
    API_BEGIN
    wavefront_check(w, n_bounces, sizes, feat, feat_stride, u, wi);
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    dfg_wavefront_tuning win{4, 1, 1, 1e30f};
    for (int pre = 0; pre <= 1; ++pre) {
        for (int bpc : {2, 4, 6, 8, 12}) {
            for (int tpf : {1, 2, 4}) {
                w->tune = dfg_wavefront_tuning{bpc, tpf, pre, 0.0f};
                wavefront_run(w, st, n_bounces, sizes, feat, feat_stride, u, guided, wi);  // capture + warm
                HIPX(hipEventRecord(e0, st));
                for (int r = 0; r < 3; ++r) wavefront_run(w, st, n_bounces, sizes, feat, feat_stride, u, guided, wi);
                HIPX(hipEventRecord(e1, st));
                HIPX(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPX(hipEventElapsedTime(&ms, e0, e1));
                ms /= 3.0f;
                if (ms < win.ms) win = dfg_wavefront_tuning{bpc, tpf, pre, ms};
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    w->tune = win;
    if (best) *best = win;
    API_END(ctx)
}

}  // extern "C"

// ============================================================================
// Request server
// ============================================================================
namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
    dfg_request r;
    double t_submit;  // seconds since the server's time origin
};

struct TrainJob {
    int64_t B;
    const float* feat;
    int stride;
    const float* target;
    const float* nrc_feat;
    const float* nrc_target;
};

struct Launch {
    std::vector<Pending> reqs;
    int64_t rows = 0;
    int64_t version = 0;
    hipEvent_t done = nullptr;
    int slot = 0;
};

}  // namespace

struct dfg_server {
    dfg_ctx* ctx = nullptr;
    dfg_server_config cfg{};
    int64_t budget_rows = 0;
    std::vector<hipStream_t> streams;
    hipStream_t train_stream = nullptr;
    hipEvent_t origin = nullptr;       // GPU time origin
    Clock::time_point t0;              // host time origin (recorded right after `origin`)
    DevBuf<WorkList> wls;              // one work list per launch slot
    DevBuf<unsigned int> counters;
    WorkList* host_wl = nullptr;       // pinned staging, one per slot
    std::vector<bool> slot_busy;
    std::deque<Pending> queue;
    std::deque<TrainJob> train_queue;
    std::deque<Launch> inflight;
    int64_t rows_inflight = 0;
    int64_t launches = 0, train_steps = 0, submitted = 0, completed = 0, rows_done = 0;
    int next_stream = 0;
    int64_t since_train = 0;
    hipEvent_t last_train = nullptr;   // queries launched after a train step wait on it
    std::vector<double> latencies;
    std::vector<int64_t> done_ids, done_versions;
    std::vector<double> done_lat;
    double first_submit = -1.0, last_done = 0.0;
    bool nrc = false;
    int nslots = 0;
    // per launch slot: compact positions and the paired grid pre-pass output
    std::vector<DevBuf<float>> slot_pos, slot_gth, slot_gph;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    ~dfg_server() {
        for (auto& l : inflight) (void)hipEventSynchronize(l.done);
        for (auto& l : inflight) (void)hipEventDestroy(l.done);
        for (auto s : streams) (void)hipStreamDestroy(s);
        if (train_stream) (void)hipStreamDestroy(train_stream);
        if (origin) (void)hipEventDestroy(origin);
        if (last_train) (void)hipEventDestroy(last_train);
        if (host_wl) (void)hipHostFree(host_wl);
        wls.release();
        counters.release();
        for (auto& b : slot_pos) b.release();
        for (auto& b : slot_gth) b.release();
        for (auto& b : slot_gph) b.release();
    }
};

namespace {

constexpr int64_t kRowBytes = 9 * 4 + 16 + 16 + 12;  // feat(>=9 floats) + u + wi + rgb per row

void server_retire(dfg_server* s, bool block) {
    while (!s->inflight.empty()) {
        Launch& l = s->inflight.front();
        if (block) HIPX(hipEventSynchronize(l.done));
        const hipError_t q = hipEventQuery(l.done);
        if (q == hipErrorNotReady) return;
        if (q != hipSuccess) throw Error(DFG_ERR_HIP, std::string("launch failed: ") + hipGetErrorString(q));
        float ms = 0.0f;
        HIPX(hipEventElapsedTime(&ms, s->origin, l.done));
        const double t_done = ms * 1e-3;
        for (auto& p : l.reqs) {
            const double lat = (t_done - p.t_submit) * 1e3;
            s->latencies.push_back(lat);
            s->done_ids.push_back(p.r.id);
            s->done_lat.push_back(lat);
            s->done_versions.push_back(l.version);
            ++s->completed;
            s->rows_done += p.r.n;
        }
        s->last_done = std::max(s->last_done, t_done);
        s->rows_inflight -= l.rows;
        s->slot_busy[l.slot] = false;
        (void)hipEventDestroy(l.done);
        s->inflight.pop_front();
    }
}

void server_train(dfg_server* s) {
    if (s->train_queue.empty()) return;
    TrainJob j = s->train_queue.front();
    s->train_queue.pop_front();
    // ordered after every launch before it ...
    for (auto& l : s->inflight) HIPX(hipStreamWaitEvent(s->train_stream, l.done, 0));
    int rc = dfg_train_step(s->ctx, s->train_stream, j.B, j.feat, j.stride, j.target, nullptr);
    if (rc != DFG_OK) throw Error(rc, s->ctx->err);
    if (j.nrc_feat && s->ctx->net[2].present) {
        rc = dfg_nrc_train_step(s->ctx, s->train_stream, j.B, j.nrc_feat, 15, j.nrc_target, nullptr);
        if (rc != DFG_OK) throw Error(rc, s->ctx->err);
    }
    // ... and before every launch after it
    HIPX(hipEventRecord(s->last_train, s->train_stream));
    ++s->train_steps;
}

bool server_launch(dfg_server* s) {
    if (s->queue.empty()) return false;
    int slot = -1;
    for (int i = 0; i < s->nslots; ++i)
        if (!s->slot_busy[i]) { slot = i; break; }
    if (slot < 0) return false;
    Launch l;
    l.slot = slot;
    WorkList& wl = s->host_wl[slot];
    wl.n_seg = 0;
    int64_t tiles = 0;
    while (!s->queue.empty() && (int)l.reqs.size() < s->cfg.max_batch_requests) {
        const Pending& p = s->queue.front();
        const bool fits = s->rows_inflight + l.rows + p.r.n <= s->budget_rows;
        if (!fits && !(l.reqs.empty() && s->inflight.empty())) break;  // an oversize request runs alone
        WorkSeg& g = wl.seg[wl.n_seg++];
        g.feat = p.r.feat;
        g.stride = p.r.feat_stride;
        g.u = p.r.u;
        g.guided = p.r.guided;
        g.wi = p.r.wi;
        g.rgb = p.r.nrc_rgb;
        g.n = p.r.n;
        g.tile_begin = tiles;
        g.gbase = l.rows;  // row of the segment in the launch's pre-pass buffers
        tiles += (p.r.n + 63) / 64;
        l.rows += p.r.n;
        l.reqs.push_back(p);
        s->queue.pop_front();
    }
    if (l.reqs.empty()) return false;
    wl.total_tiles = tiles;
    wl.total_rows = l.rows;
    const int64_t n_reqs = (int64_t)l.reqs.size();
    hipStream_t st = s->streams[s->next_stream];
    s->next_stream = (s->next_stream + 1) % (int)s->streams.size();
    if (s->train_steps > 0) HIPX(hipStreamWaitEvent(st, s->last_train, 0));
    WorkList* dwl = s->wls.p + slot;
    const size_t bytes = offsetof(WorkList, seg) + sizeof(WorkSeg) * (size_t)wl.n_seg;
    HIPX(hipMemcpyAsync(dwl, &wl, bytes, hipMemcpyHostToDevice, st));
    // counter reset by a fill kernel (graph-safe), then the persistent launch
    unsigned int* ctr = s->counters.p + slot;
    launch_fill_u32(ctr, 0u, 1, st);
    QueryArgs a = base_args(s->ctx);
    a.feat = l.reqs[0].r.feat;
    a.stride = l.reqs[0].r.feat_stride;
    a.n = l.rows;
    a.rgb = s->nrc ? l.reqs[0].r.nrc_rgb : nullptr;
    dfg_ctx* c = s->ctx;
    if (c->pair_ok && l.rows >= c->prepass_min_rows) {
        // the XCD-aware paired grid pre-pass over every row of the launch
        const int64_t ld = round_up(l.rows, 64);
        const size_t cols = (size_t)c->net[0].spec.grid.n_levels * 2;
        s->slot_pos[slot].ensure((size_t)l.rows * 3);
        s->slot_gth[slot].ensure(cols * ld);
        s->slot_gph[slot].ensure(cols * ld);
        launch_extract_pos_worklist(dwl, grid_pos_dim(c->net[0].spec), l.rows, s->slot_pos[slot].p, st);
        launch_grid_encode_pair(c->net[0].d_spec.p, c->net[0].spec, c->tab_pair.p, s->slot_pos[slot].p, l.rows, ld,
                                s->slot_gth[slot].p, s->slot_gph[slot].p, st);
        a.gfeat_th = s->slot_gth[slot].p;
        a.gfeat_ph = s->slot_gph[slot].p;
        a.gld = ld;
    }
    fill_lds_strides(c, a, s->nrc, true);
    const int64_t want_blocks = std::min<int64_t>(tiles, (int64_t)cu_count() * 4);
    launch_query_worklist(a, s->ctx->n_theta, s->ctx->n_phi, dwl, ctr, (int)std::max<int64_t>(1, want_blocks), 1, st);
    check_launch();
    HIPX(hipEventCreateWithFlags(&l.done, 0));
    HIPX(hipEventRecord(l.done, st));
    l.version = s->train_steps;
    s->rows_inflight += l.rows;
    s->slot_busy[slot] = true;
    s->inflight.push_back(std::move(l));
    ++s->launches;
    // a training step is due after every `train_every` query batches; it is
    // enqueued right behind the launch that completes the count
    s->since_train += n_reqs;
    while (s->cfg.train_every > 0 && s->since_train >= s->cfg.train_every && !s->train_queue.empty()) {
        server_train(s);
        s->since_train -= s->cfg.train_every;
    }
    return true;
}

}  // namespace

extern "C" {

int dfg_server_create(dfg_ctx* ctx, const dfg_server_config* cfg, dfg_server** out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!cfg || !out) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (cfg->budget_bytes < kRowBytes * 64) throw Error(DFG_ERR_BAD_SIZE, "budget below one tile of rows");
    if (cfg->max_batch_requests < 1 || cfg->max_batch_requests > DFG_MAX_SEGS)
        throw Error(DFG_ERR_BAD_SIZE, "max_batch_requests must be 1..256");
    if (cfg->n_streams < 1 || cfg->n_streams > 4) throw Error(DFG_ERR_BAD_SIZE, "n_streams must be 1..4");
    if (cfg->train_every < 0) throw Error(DFG_ERR_BAD_SIZE, "train_every < 0");
    std::unique_ptr<dfg_server> s(new dfg_server());
    s->ctx = ctx;
    s->cfg = *cfg;
    s->budget_rows = cfg->budget_bytes / kRowBytes;
    for (int i = 0; i < cfg->n_streams; ++i) {
        hipStream_t st;
        HIPX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->streams.push_back(st);
    }
    HIPX(hipStreamCreateWithFlags(&s->train_stream, hipStreamNonBlocking));
    s->nslots = cfg->n_streams * 2;
    s->slot_pos.resize(s->nslots);
    s->slot_gth.resize(s->nslots);
    s->slot_gph.resize(s->nslots);
    s->wls.ensure(s->nslots);
    s->counters.ensure(s->nslots);
    HIPX(hipHostMalloc((void**)&s->host_wl, sizeof(WorkList) * s->nslots, hipHostMallocDefault));
    s->slot_busy.assign(s->nslots, false);
    HIPX(hipEventCreate(&s->last_train));
    HIPX(hipEventCreate(&s->origin));
    HIPX(hipDeviceSynchronize());
    HIPX(hipEventRecord(s->origin, s->streams[0]));
    HIPX(hipEventSynchronize(s->origin));
    s->t0 = Clock::now();
    *out = s.release();
    API_END(ctx)
}

int dfg_server_destroy(dfg_server* s) {
    if (!s) return DFG_ERR_NOT_INIT;
    delete s;
    return DFG_OK;
}

int dfg_server_submit(dfg_server* s, const dfg_request* r) {
    if (!s) return DFG_ERR_NOT_INIT;
    dfg_ctx* ctx = s->ctx;
    API_BEGIN
    if (!r) throw Error(DFG_ERR_INVALID_ARG, "null request");
    if (r->n < 1 || r->n > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "request size out of range");
    if (!r->feat || !r->u || !r->wi) throw Error(DFG_ERR_INVALID_ARG, "null request buffer");
    if (r->feat_stride < ctx->D) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below n_input_dims");
    const bool wants_nrc = r->nrc_rgb != nullptr;
    if (wants_nrc && (!ctx->net[2].present || r->feat_stride < 15))
        throw Error(DFG_ERR_INVALID_ARG, "NRC output needs an NRC net and feat_stride >= 15");
    if (s->submitted > 0 && wants_nrc != s->nrc) throw Error(DFG_ERR_INVALID_ARG, "mixed NRC and non-NRC requests");
    s->nrc = wants_nrc;
    const double t = s->now();
    if (s->first_submit < 0) s->first_submit = t;
    s->queue.push_back(Pending{*r, t});
    ++s->submitted;
    API_END(ctx)
}

int dfg_server_submit_train(dfg_server* s, int64_t B, const float* feat, int feat_stride, const float* target,
                            const float* nrc_feat, const float* nrc_target) {
    if (!s) return DFG_ERR_NOT_INIT;
    dfg_ctx* ctx = s->ctx;
    API_BEGIN
    if (B < 1 || !feat || !target) throw Error(DFG_ERR_INVALID_ARG, "bad training job");
    if (nrc_feat && !nrc_target) throw Error(DFG_ERR_INVALID_ARG, "NRC job needs targets");
    s->train_queue.push_back(TrainJob{B, feat, feat_stride, target, nrc_feat, nrc_target});
    API_END(ctx)
}

int dfg_server_progress(dfg_server* s, int wait) {
    if (!s) return DFG_ERR_NOT_INIT;
    dfg_ctx* ctx = s->ctx;
    API_BEGIN
    for (;;) {
        server_retire(s, false);
        while (server_launch(s)) {
        }
        if (!wait) break;
        if (s->queue.empty() && s->inflight.empty()) {
            if (s->train_steps > 0) {  // the busy span ends with the last training step too
                HIPX(hipEventSynchronize(s->last_train));
                float ms = 0.0f;
                HIPX(hipEventElapsedTime(&ms, s->origin, s->last_train));
                s->last_done = std::max(s->last_done, ms * 1e-3);
            }
            break;
        }
        if (!s->inflight.empty()) server_retire(s, true);
    }
    API_END(ctx)
}

int dfg_server_get_stats(dfg_server* s, dfg_server_stats* out) {
    if (!s) return DFG_ERR_NOT_INIT;
    if (!out) return DFG_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->launches = s->launches;
    out->train_steps = s->train_steps;
    out->rows = s->rows_done;
    if (!s->latencies.empty()) {
        std::vector<double> v = s->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double idx = p * (double)(v.size() - 1);
            const size_t lo = (size_t)std::floor(idx), hi = std::min(lo + 1, v.size() - 1);
            return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        out->mean_ms = sum / (double)v.size();
        out->max_ms = v.back();
    }
    out->busy_s = s->first_submit >= 0 ? std::max(0.0, s->last_done - s->first_submit) : 0.0;
    return DFG_OK;
}

int dfg_server_drain(dfg_server* s, int64_t* ids, double* latency_ms, int64_t* weights_version, int max, int* count) {
    if (!s) return DFG_ERR_NOT_INIT;
    if (!count || max < 0) return DFG_ERR_INVALID_ARG;
    const int k = (int)std::min<size_t>((size_t)max, s->done_ids.size());
    for (int i = 0; i < k; ++i) {
        if (ids) ids[i] = s->done_ids[i];
        if (latency_ms) latency_ms[i] = s->done_lat[i];
        if (weights_version) weights_version[i] = s->done_versions[i];
    }
    s->done_ids.erase(s->done_ids.begin(), s->done_ids.begin() + k);
    s->done_lat.erase(s->done_lat.begin(), s->done_lat.begin() + k);
    s->done_versions.erase(s->done_versions.begin(), s->done_versions.begin() + k);
    *count = k;
    return DFG_OK;
}

}  // extern "C"
