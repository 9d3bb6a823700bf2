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
#include <thread>
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
    fill_lds_strides(c, a, false, true, 1);
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
// Request server.  Admission runs on the device (k_serve.hip): the host pushes
// requests into a ring in pinned memory and launches rounds; the first kernel
// of each round takes what the row budget allows from the ring's head and
// builds the round's work list there.  The host's part is the ring, launching
// rounds, the training interleave and the latency records.
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

struct Round {
    int slot = 0;
    uint32_t seq = 0;        // the slot's admission count this round publishes
    bool published = false;
    uint32_t h0 = 0, h1 = 0;  // ring positions the round admitted
    int64_t rows = 0;
    int64_t version = 0;      // training steps enqueued before the round
    hipEvent_t done = nullptr;
};

// One round in flight at a time per slot; slot i launches on stream i % n_streams.
struct Slot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t launched = 0;
    bool busy = false;
    DevBuf<float> pos, gth, gph;  // the round's positions and paired pre-pass output
};

constexpr int64_t kRowBytes = 9 * 4 + 16 + 16 + 12;  // feat(>=9 floats) + u + wi + rgb per row
// per slot: the query's tile counter, then 8 pre-pass item counters, each on a
// 128-byte line of its own
constexpr size_t kSlotCounters = 32 + 32 * kXcdGroups;

}  // namespace

struct dfg_server {
    dfg_ctx* ctx = nullptr;
    dfg_server_config cfg{};
    int64_t budget_rows = 0;
    int64_t round_rows = 0;
    std::vector<hipStream_t> streams;
    hipStream_t train_stream = nullptr;
    hipStream_t cap_stream = nullptr;
    hipEvent_t origin = nullptr;       // GPU time origin
    Clock::time_point t0;              // host time origin (recorded right after `origin`)
    ServerHostShared* hs = nullptr;    // pinned: ring, tail, admission records (host address)
    ServerHostShared* hs_dev = nullptr;  // the same memory's device address
    DevBuf<ServerDevState> dv;
    DevBuf<WorkList> wls;              // one work list per slot
    DevBuf<unsigned int> counters;     // per slot: the query's tile counter, then the pre-pass's per-XCD item counters
    std::vector<Slot> slots;
    std::deque<Pending> pend;          // submitted, not completed; pend[i] is ring position pend_base + i
    uint32_t pend_base = 0;
    uint32_t pushed = 0;               // ring positions written (the tail)
    uint32_t known_head = 0;           // the head after the newest published admission
    std::deque<TrainJob> train_queue;
    std::deque<Round> inflight;        // in launch order
    bool unpublished = false;          // the newest round has not published its admission yet
    bool blocked = false;              // the newest admission took nothing while requests waited
    int64_t launches = 0, rounds = 0, train_steps = 0, submitted = 0, completed = 0, rows_done = 0;
    int64_t since_train = 0;
    hipEvent_t last_train = nullptr;   // rounds launched after a train step wait on it
    std::vector<double> latencies;
    std::vector<int64_t> done_ids, done_versions;
    std::vector<double> done_lat;
    double first_submit = -1.0, last_done = 0.0;
    bool nrc = false;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    ~dfg_server() {
        for (auto& l : inflight) (void)hipEventSynchronize(l.done);
        for (auto& l : inflight) (void)hipEventDestroy(l.done);
        if (last_train) (void)hipEventSynchronize(last_train);
        for (auto& sl : slots) {
            if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
            if (sl.graph) (void)hipGraphDestroy(sl.graph);
            sl.pos.release();
            sl.gth.release();
            sl.gph.release();
        }
        for (auto s : streams) (void)hipStreamDestroy(s);
        if (train_stream) (void)hipStreamDestroy(train_stream);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (origin) (void)hipEventDestroy(origin);
        if (last_train) (void)hipEventDestroy(last_train);
        if (hs) (void)hipHostFree(hs);
        dv.release();
        wls.release();
        counters.release();
    }
};

namespace {

uint32_t load_acquire(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }

// Writes queued requests into the ring while it has room: positions below the
// newest published head have been copied into a work list and are free.
void server_push(dfg_server* s) {
    const uint32_t end = s->pend_base + (uint32_t)s->pend.size();
    bool any = false;
    while (s->pushed != end && s->pushed - s->known_head < kServerRing) {
        const dfg_request& r = s->pend[s->pushed - s->pend_base].r;
        ServerReq& q = s->hs->ring[s->pushed & (kServerRing - 1)];
        q.id = r.id;
        q.n = r.n;
        q.feat = r.feat;
        q.u = r.u;
        q.guided = r.guided;
        q.wi = r.wi;
        q.rgb = r.nrc_rgb;
        q.stride = r.feat_stride;
        ++s->pushed;
        any = true;
    }
    if (any) __atomic_store_n(&s->hs->tail, s->pushed, __ATOMIC_RELEASE);
}

void server_train(dfg_server* s) {
    if (s->train_queue.empty()) return;
    TrainJob j = s->train_queue.front();
    s->train_queue.pop_front();
    // ordered after every round before it ...
    for (auto& l : s->inflight) HIPX(hipStreamWaitEvent(s->train_stream, l.done, 0));
    int rc = dfg_train_step(s->ctx, s->train_stream, j.B, j.feat, j.stride, j.target, nullptr);
    if (rc != DFG_OK) throw Error(rc, s->ctx->err);
    if (j.nrc_feat && s->ctx->net[2].present) {
        rc = dfg_nrc_train_step(s->ctx, s->train_stream, j.B, j.nrc_feat, 15, j.nrc_target, nullptr);
        if (rc != DFG_OK) throw Error(rc, s->ctx->err);
    }
    // ... and before every round after it
    HIPX(hipEventRecord(s->last_train, s->train_stream));
    ++s->train_steps;
}

// Reads the newest round's admission record once the device has published it.
void server_poll(dfg_server* s) {
    if (!s->unpublished) return;
    Round& l = s->inflight.back();
    const ServerRecord& rec = s->hs->rec[l.slot];
    if (load_acquire(&rec.seq) != l.seq) return;
    l.h0 = rec.h0;
    l.h1 = rec.h1;
    l.rows = rec.rows;
    l.published = true;
    s->unpublished = false;
    if (l.h0 != s->known_head) throw Error(DFG_ERR_HIP, "server admission out of order");
    s->known_head = l.h1;
    const uint32_t k = l.h1 - l.h0;
    s->blocked = k == 0;
    if (k) {
        ++s->rounds;
        // a training step is due after every `train_every` query batches; it
        // is enqueued right behind the round that completes the count
        s->since_train += k;
        while (s->cfg.train_every > 0 && s->since_train >= s->cfg.train_every && !s->train_queue.empty()) {
            server_train(s);
            s->since_train -= s->cfg.train_every;
        }
    }
    server_push(s);
}

// Waits (bounded) until the newest round has published its admission.
void server_wait_published(dfg_server* s) {
    const auto limit = Clock::now() + std::chrono::seconds(30);
    while (s->unpublished) {
        server_poll(s);
        if (!s->unpublished) break;
        if (Clock::now() > limit) throw Error(DFG_ERR_HIP, "server round did not publish its admission");
        std::this_thread::yield();
    }
}

void server_retire(dfg_server* s, bool block) {
    while (!s->inflight.empty()) {
        Round& l = s->inflight.front();
        if (block) HIPX(hipEventSynchronize(l.done));
        const hipError_t q = hipEventQuery(l.done);
        if (q == hipErrorNotReady) return;
        if (q != hipSuccess) throw Error(DFG_ERR_HIP, std::string("round failed: ") + hipGetErrorString(q));
        if (!l.published) server_poll(s);  // finished rounds have published (their first kernel did)
        if (!l.published) throw Error(DFG_ERR_HIP, "finished round without an admission record");
        float ms = 0.0f;
        HIPX(hipEventElapsedTime(&ms, s->origin, l.done));
        const double t_done = ms * 1e-3;
        for (uint32_t pos = l.h0; pos != l.h1; ++pos) {
            if (s->pend.empty() || s->pend_base != pos) throw Error(DFG_ERR_HIP, "server retired out of order");
            const Pending& p = s->pend.front();
            const double lat = (t_done - p.t_submit) * 1e3;
            s->latencies.push_back(lat);
            s->done_ids.push_back(p.r.id);
            s->done_lat.push_back(lat);
            s->done_versions.push_back(l.version);
            ++s->completed;
            s->rows_done += p.r.n;
            s->pend.pop_front();
            ++s->pend_base;
        }
        if (l.h1 != l.h0) s->last_done = std::max(s->last_done, t_done);
        s->slots[l.slot].busy = false;
        s->blocked = false;  // rows went back to the budget
        (void)hipEventDestroy(l.done);
        s->inflight.pop_front();
    }
}

// Captures the slot's round: admission, positions, paired pre-pass, the
// persistent work-list query and the completion, all sized on the device.
void server_capture(dfg_server* s, int slot) {
    dfg_ctx* c = s->ctx;
    Slot& sl = s->slots[slot];
    WorkList* dwl = s->wls.p + slot;
    unsigned int* ctr = s->counters.p + (size_t)slot * kSlotCounters;
    unsigned int* item_ctr = ctr + 32;
    const int64_t R = s->round_rows;
    const int64_t ld = round_up(R, 64);
    QueryArgs a = base_args(c);
    a.n = R;
    a.rgb = s->nrc ? s->pend.front().r.nrc_rgb : nullptr;  // selects the NRC variant; rows come from the work list
    if (c->pair_ok) {
        const size_t cols = (size_t)c->net[0].spec.grid.n_levels * 2;
        sl.pos.ensure((size_t)R * 3);
        sl.gth.ensure(cols * ld);
        sl.gph.ensure(cols * ld);
        a.gfeat_th = sl.gth.p;
        a.gfeat_ph = sl.gph.p;
        a.gld = ld;
    }
    fill_lds_strides(c, a, s->nrc, true, 1);
    AdmitParams p{s->budget_rows, R, s->cfg.max_batch_requests, slot};
    hipStream_t st = s->cap_stream;
    HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_server_admit(s->hs_dev, s->dv.p, dwl, ctr, item_ctr, p, st);
    if (c->pair_ok) {
        launch_extract_pos_round(dwl, grid_pos_dim(c->net[0].spec), R, sl.pos.p, st);
        launch_grid_encode_pair_dev(c->net[0].d_spec.p, c->net[0].spec, c->tab_pair.p, sl.pos.p, R, ld, sl.gth.p,
                                    sl.gph.p, &dwl->total_rows, item_ctr, cu_count() * 8, st);
    }
    // 8 workgroups (waves) per CU pulling 2 tiles at a time: the autotuner's
    // choice for full-size frames of the same work-list kernel
    launch_query_worklist(a, c->n_theta, c->n_phi, dwl, ctr, cu_count() * 8, 2, st);
    launch_server_complete(s->dv.p, dwl, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) throw Error(DFG_ERR_HIP, std::string("server graph capture: ") + hipGetErrorString(e));
    sl.graph = g;
    HIPX(hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0));
}

bool server_launch(dfg_server* s) {
    // with nothing in flight the whole budget is free and a round admits at
    // least its first request
    if (s->inflight.empty()) s->blocked = false;
    if (s->unpublished || s->blocked) return false;
    server_push(s);
    if (s->pushed == s->known_head) return false;  // nothing waiting in the ring
    int slot = -1;
    for (int i = 0; i < (int)s->slots.size(); ++i)
        if (!s->slots[i].busy) { slot = i; break; }
    if (slot < 0) return false;
    Slot& sl = s->slots[slot];
    if (!sl.exec) server_capture(s, slot);
    hipStream_t st = s->streams[slot % s->streams.size()];
    if (s->train_steps > 0) HIPX(hipStreamWaitEvent(st, s->last_train, 0));
    HIPX(hipGraphLaunch(sl.exec, st));
    Round l;
    l.slot = slot;
    l.seq = ++sl.launched;
    l.version = s->train_steps;
    HIPX(hipEventCreateWithFlags(&l.done, 0));
    HIPX(hipEventRecord(l.done, st));
    sl.busy = true;
    s->unpublished = true;
    s->inflight.push_back(l);
    ++s->launches;
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
    if (cfg->round_rows < 0 || cfg->round_rows > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "round_rows out of range");
    std::unique_ptr<dfg_server> s(new dfg_server());
    s->ctx = ctx;
    s->cfg = *cfg;
    s->budget_rows = std::min<int64_t>(cfg->budget_bytes / kRowBytes, ((int64_t)1 << 31) - 1);
    s->round_rows = cfg->round_rows > 0 ? cfg->round_rows
                                        : std::max<int64_t>((int64_t)1 << 20, std::min<int64_t>(s->budget_rows, (int64_t)1 << 22));
    for (int i = 0; i < cfg->n_streams; ++i) {
        hipStream_t st;
        HIPX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->streams.push_back(st);
    }
    HIPX(hipStreamCreateWithFlags(&s->train_stream, hipStreamNonBlocking));
    HIPX(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
    s->slots.resize((size_t)cfg->n_streams * 2);
    s->wls.ensure(s->slots.size());
    s->counters.ensure(s->slots.size() * kSlotCounters);
    s->dv.ensure(1);
    HIPX(hipMemset(s->dv.p, 0, sizeof(ServerDevState)));
    HIPX(hipHostMalloc((void**)&s->hs, sizeof(ServerHostShared), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(s->hs, 0, sizeof(ServerHostShared));
    HIPX(hipHostGetDevicePointer((void**)&s->hs_dev, s->hs, 0));
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
    if (r->n < 1 || r->n > s->round_rows) throw Error(DFG_ERR_BAD_SIZE, "request size out of range (1..round_rows)");
    if (!r->feat || !r->u || !r->wi) throw Error(DFG_ERR_INVALID_ARG, "null request buffer");
    if (r->feat_stride < ctx->D) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below n_input_dims");
    const bool wants_nrc = r->nrc_rgb != nullptr;
    if (wants_nrc && (!ctx->net[2].present || r->feat_stride < 15))
        throw Error(DFG_ERR_INVALID_ARG, "NRC output needs an NRC net and feat_stride >= 15");
    if (s->submitted > 0 && wants_nrc != s->nrc) throw Error(DFG_ERR_INVALID_ARG, "mixed NRC and non-NRC requests");
    s->nrc = wants_nrc;
    const double t = s->now();
    if (s->first_submit < 0) s->first_submit = t;
    s->pend.push_back(Pending{*r, t});
    ++s->submitted;
    // visible to a round whose admission runs from now on
    server_push(s);
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
        server_poll(s);
        server_retire(s, false);
        if (server_launch(s)) continue;
        if (!wait) break;
        if (s->pend.empty() && s->inflight.empty()) {
            if (s->train_steps > 0) {  // the busy span ends with the last training step too
                HIPX(hipEventSynchronize(s->last_train));
                float ms = 0.0f;
                HIPX(hipEventElapsedTime(&ms, s->origin, s->last_train));
                s->last_done = std::max(s->last_done, ms * 1e-3);
            }
            break;
        }
        if (s->unpublished) {
            server_wait_published(s);
        } else if (!s->inflight.empty()) {
            server_retire(s, true);
        } else {
            break;  // unreachable: with nothing in flight server_launch takes the next request
        }
    }
    API_END(ctx)
}

int dfg_server_get_stats(dfg_server* s, dfg_server_stats* out) {
    if (!s) return DFG_ERR_NOT_INIT;
    if (!out) return DFG_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->launches = s->rounds;
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
    unsigned long long peak = 0;
    if (hipMemcpy(&peak, &s->dv.p->peak_rows, sizeof(peak), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(s->ctx, DFG_ERR_HIP, "server stats: device read failed");
    out->peak_rows = (int64_t)peak;
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
