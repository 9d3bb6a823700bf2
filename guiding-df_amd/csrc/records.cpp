// Training from recorded path vertices: row generation entry points and the
// reference's per-frame training passes (LearnedDistribution::trainStep,
// learnedDistribution.cpp:208-293; NRC::trainStep, nrc.cpp:114-170).
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.h"

using namespace dfg;

namespace dfg {
void launch_training_targets(const dfg_records& r, int64_t n_train, uint32_t offset, uint32_t stride,
                             const float* aabb, int mode, float* feat, float* tgt, int2* vidx, int64_t capacity,
                             uint32_t* local, int64_t* block_tot, int64_t* count, hipStream_t s);
void launch_nrc_substitute(const dfg_records& r, const int2* vidx, const float* rgb, const int64_t* count,
                           int64_t capacity, hipStream_t s);
void launch_directions(int64_t n, const float* wi, float* dir, float* pdf, hipStream_t s);
void launch_pdf_image_rows(const float* feat, int D, int rt, int rp, float* rows, float* wi2, hipStream_t s);
void launch_scatter(int64_t n, float f, const float* wi, const uint8_t* guided, const uint8_t* mix,
                    const float* bsdf_pdf, const float* bsdf_val, const float* cos_theta, float* thp, float* spdf,
                    float* mis, uint8_t* alive, hipStream_t s);
void launch_sample_wi(int64_t n, const SampleWiArgs& a, hipStream_t s);
void launch_record_scatter(int64_t n, const dfg_records& r, const dfg_shading& sd, const RecordScatterArgs& a,
                           hipStream_t s);
void launch_record_radiance(int64_t n, const dfg_records& r, const RecordRadianceArgs& a, hipStream_t s);
}

namespace {

constexpr int64_t kMinTrainBatch = 65536;  // MIN_TRAIN_BATCH_SIZE, parameters.h:55

void check_records(const dfg_records* r, int64_t n_train, uint32_t offset, uint32_t stride, const float* aabb,
                   int mode) {
    if (!r || !aabb) throw Error(DFG_ERR_INVALID_ARG, "null records or aabb");
    if (r->max_depth < 1 || r->n_pixels < 1) throw Error(DFG_ERR_BAD_SIZE, "empty record buffer");
    if (n_train < 0) throw Error(DFG_ERR_BAD_SIZE, "n_train < 0");
    if (n_train > 0 && stride < 1) throw Error(DFG_ERR_BAD_SIZE, "pixel_stride < 1");
    if (n_train > 0 && (int64_t)offset + (n_train - 1) * (int64_t)stride >= r->n_pixels)
        throw Error(DFG_ERR_BAD_SIZE, "training pixels beyond n_pixels");
    const void* need[] = {r->cur_depth, r->pos, r->wo, r->normal, r->roughness, r->L, r->thp, r->bsdf_val,
                          r->wi_pdf, r->mis_weight, r->delta};
    for (const void* p : need)
        if (!p) throw Error(DFG_ERR_INVALID_ARG, "null record field");
    if (mode == 0 && !r->dir) throw Error(DFG_ERR_INVALID_ARG, "null record field dir");
    if (mode != 0 && (!r->diffuse || !r->specular || !r->miss)) throw Error(DFG_ERR_INVALID_ARG, "null NRC record field");
}

void targets(dfg_ctx* c, hipStream_t st, const dfg_records* r, int64_t n_train, uint32_t offset, uint32_t stride,
             const float* aabb, int mode, float* feat, float* tgt, int64_t capacity, int64_t* count) {
    if (n_train == 0) {
        launch_fill_u32((unsigned int*)count, 0u, 2, st);
        return;
    }
    const int64_t blocks = (n_train + 255) / 256;
    c->rec_local.ensure((size_t)n_train);
    c->rec_blk.ensure((size_t)blocks);
    launch_training_targets(*r, n_train, offset, stride, aabb, mode, feat, tgt, nullptr, capacity, c->rec_local.p,
                            c->rec_blk.p, count, st);
    check_launch();
}

// rows into the context's training buffer; the row count on the host
int64_t targets_to_buffer(dfg_ctx* c, hipStream_t st, const dfg_records* r, int64_t n_train, uint32_t offset,
                          uint32_t stride, const float* aabb, int mode) {
    const int64_t cap = n_train * r->max_depth;
    const int nf = mode == 0 ? 9 : 15, nt = mode == 0 ? 9 : 3;
    c->rec_feat.ensure((size_t)std::max<int64_t>(cap, 1) * nf);
    c->rec_tgt.ensure((size_t)std::max<int64_t>(cap, 1) * nt);
    c->rec_count.ensure(1);
    targets(c, st, r, n_train, offset, stride, aabb, mode, c->rec_feat.p, c->rec_tgt.p, cap, c->rec_count.p);
    int64_t rows = 0;
    HIPX(hipMemcpyAsync(&rows, c->rec_count.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    return std::min(rows, cap);
}

}  // namespace

extern "C" {

int dfg_training_targets(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train, uint32_t pixel_offset,
                         uint32_t pixel_stride, const float* aabb, float* feat, float* target, int64_t capacity,
                         int64_t* count) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_records(rec, n_train, pixel_offset, pixel_stride, aabb, 0);
    if (!count || capacity < 0 || (capacity > 0 && (!feat || !target))) throw Error(DFG_ERR_INVALID_ARG, "bad output");
    targets(ctx, (hipStream_t)stream, rec, n_train, pixel_offset, pixel_stride, aabb, 0, feat, target, capacity, count);
    API_END(ctx)
}

int dfg_nrc_training_targets(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train,
                             uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb, float* feat,
                             float* target_rgb, int64_t capacity, int64_t* count) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_records(rec, n_train, pixel_offset, pixel_stride, aabb, 1);
    if (!count || capacity < 0 || (capacity > 0 && (!feat || !target_rgb)))
        throw Error(DFG_ERR_INVALID_ARG, "bad output");
    targets(ctx, (hipStream_t)stream, rec, n_train, pixel_offset, pixel_stride, aabb, 1, feat, target_rgb, capacity,
            count);
    API_END(ctx)
}

int dfg_train_records(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train, uint32_t pixel_offset,
                      uint32_t pixel_stride, const float* aabb, int64_t batch_size, int batch_per_frame,
                      int64_t* n_rows, int* n_steps, float* loss) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_records(rec, n_train, pixel_offset, pixel_stride, aabb, 0);
    if (batch_size < 128 || batch_per_frame < 1) throw Error(DFG_ERR_BAD_SIZE, "bad batch schedule");
    hipStream_t st = (hipStream_t)stream;
    const int64_t rows = targets_to_buffer(ctx, st, rec, n_train, pixel_offset, pixel_stride, aabb, 0);
    // learnedDistribution.cpp:226-236
    const int planned = (int)std::min<int64_t>(rows / batch_size + 1, batch_per_frame);
    ctx->rec_loss.ensure((size_t)planned);
    int run = 0;
    for (int it = 0; it < planned; ++it) {
        int64_t b = std::min<int64_t>(rows - (int64_t)it * batch_size, batch_size);
        b -= b % 128;
        if (b < kMinTrainBatch) break;
        const int64_t r0 = (int64_t)it * batch_size;
        const int rc = dfg_train_step(ctx, st, b, ctx->rec_feat.p + r0 * 9, 9, ctx->rec_tgt.p + r0 * 9,
                                      ctx->rec_loss.p + it);
        if (rc != DFG_OK) throw Error(rc, ctx->err);
        ++run;
    }
    std::vector<float> l((size_t)std::max(run, 1), 0.0f);
    if (run) HIPX(hipMemcpyAsync(l.data(), ctx->rec_loss.p, sizeof(float) * run, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    float sum = 0.0f;
    for (int i = 0; i < run; ++i) sum += l[i];
    if (n_rows) *n_rows = rows;
    if (n_steps) *n_steps = run;
    if (loss) *loss = sum / (float)planned;
    API_END(ctx)
}

int dfg_nrc_train_records(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train, uint32_t pixel_offset,
                          uint32_t pixel_stride, const float* aabb, int64_t batch_size, int batch_per_frame,
                          int64_t* n_rows, int* n_steps, float* loss) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    check_records(rec, n_train, pixel_offset, pixel_stride, aabb, 1);
    if (batch_size < 128 || batch_per_frame < 1) throw Error(DFG_ERR_BAD_SIZE, "bad batch schedule");
    hipStream_t st = (hipStream_t)stream;
    const int64_t rows = targets_to_buffer(ctx, st, rec, n_train, pixel_offset, pixel_stride, aabb, 1);
    // nrc.cpp:127-138, including its slice offset (step k starts at k times its size)
    const int planned = (int)std::max<int64_t>(rows / batch_size + 1, batch_per_frame);
    const int64_t eff = std::min<int64_t>(rows / planned, batch_size);
    ctx->rec_loss.ensure((size_t)planned);
    int run = 0;
    for (int it = 0; it < planned; ++it) {
        int64_t b = std::min<int64_t>(rows - (int64_t)it * eff, eff);
        b -= b % 128;
        if (b < kMinTrainBatch) break;
        const int64_t r0 = (int64_t)it * b;
        const int rc = dfg_nrc_train_step(ctx, st, b, ctx->rec_feat.p + r0 * 15, 15, ctx->rec_tgt.p + r0 * 3,
                                          ctx->rec_loss.p + it);
        if (rc != DFG_OK) throw Error(rc, ctx->err);
        ++run;
    }
    std::vector<float> l((size_t)std::max(run, 1), 0.0f);
    if (run) HIPX(hipMemcpyAsync(l.data(), ctx->rec_loss.p, sizeof(float) * run, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    float sum = 0.0f;
    for (int i = 0; i < run; ++i) sum += l[i];
    if (n_rows) *n_rows = rows;
    if (n_steps) *n_steps = run;
    if (loss) *loss = sum / (float)planned;
    API_END(ctx)
}

int dfg_nrc_substitute_radiance(dfg_ctx* ctx, void* stream, dfg_records* rec, int64_t n_train, uint32_t pixel_offset,
                                uint32_t pixel_stride, const float* aabb, int64_t* n_inferred) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    check_records(rec, n_train, pixel_offset, pixel_stride, aabb, 2);
    hipStream_t st = (hipStream_t)stream;
    int64_t m = 0;
    if (n_train > 0) {
        const int64_t cap = n_train * rec->max_depth;
        const int64_t blocks = (n_train + 255) / 256;
        ctx->rec_local.ensure((size_t)n_train);
        ctx->rec_blk.ensure((size_t)blocks);
        ctx->rec_feat.ensure((size_t)cap * 15);
        ctx->rec_tgt.ensure((size_t)cap * 3);
        ctx->rec_vidx.ensure((size_t)cap);
        ctx->rec_count.ensure(1);
        // 1. NRC inputs of the vertices below cur_depth; cur_depth lowered by one
        launch_training_targets(*rec, n_train, pixel_offset, pixel_stride, aabb, 2, ctx->rec_feat.p, nullptr,
                                ctx->rec_vidx.p, cap, ctx->rec_local.p, ctx->rec_blk.p, ctx->rec_count.p, st);
        check_launch();
        HIPX(hipMemcpyAsync(&m, ctx->rec_count.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        m = std::min(m, cap);
        // 2. the radiance cache at those vertices
        if (m > 0) {
            const int rc = dfg_nrc_infer(ctx, st, m, ctx->rec_feat.p, 15, ctx->rec_tgt.p);
            if (rc != DFG_OK) throw Error(rc, ctx->err);
            // 3. replaceIndirectRadianceSkipDelta, then divideRadiance
            launch_nrc_substitute(*rec, ctx->rec_vidx.p, ctx->rec_tgt.p, ctx->rec_count.p, m, st);
            check_launch();
        }
        HIPX(hipStreamSynchronize(st));
    }
    if (n_inferred) *n_inferred = m;
    API_END(ctx)
}

int dfg_pdf_image(dfg_ctx* ctx, void* stream, const float* feat, int res_theta, int res_phi, float* image) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!feat || !image) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (res_theta < 1 || res_phi < 1 || res_theta > 4096 || res_phi > 4096)
        throw Error(DFG_ERR_BAD_SIZE, "image resolution must be 1..4096");
    const int64_t n = (int64_t)res_theta * res_phi;
    ctx->img_rows.ensure((size_t)n * ctx->D);
    ctx->img_wi2.ensure((size_t)n * 2);
    hipStream_t st = (hipStream_t)stream;
    launch_pdf_image_rows(feat, ctx->D, res_theta, res_phi, ctx->img_rows.p, ctx->img_wi2.p, st);
    check_launch();
    const int rc = dfg_pdf(ctx, st, n, ctx->img_rows.p, ctx->D, ctx->img_wi2.p, image);
    if (rc != DFG_OK) throw Error(rc, ctx->err);
    API_END(ctx)
}

int dfg_directions(void* stream, int64_t n, const float* wi, float* dir, float* pdf) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (n == 0) return DFG_OK;
    if (!wi) return DFG_ERR_INVALID_ARG;
    launch_directions(n, wi, dir, pdf, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

int dfg_scatter(void* stream, int64_t n, float bsdf_fraction, const float* wi, const uint8_t* guided,
                const uint8_t* mix, const float* bsdf_pdf, const float* bsdf_val, const float* cos_theta,
                float* thp, float* scatter_pdf, float* mis, uint8_t* alive) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (!(bsdf_fraction >= 0.0f && bsdf_fraction <= 1.0f)) return DFG_ERR_INVALID_ARG;
    if (n == 0) return DFG_OK;
    if (!wi || !bsdf_pdf || !bsdf_val || !cos_theta || !thp) return DFG_ERR_INVALID_ARG;
    launch_scatter(n, bsdf_fraction, wi, guided, mix, bsdf_pdf, bsdf_val, cos_theta, thp, scatter_pdf, mis, alive,
                   (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

int dfg_sample_wi(void* stream, int64_t n, float rr_prob, float bsdf_fraction, int guiding_enabled,
                  int max_guided_depth, const int32_t* depth, const uint8_t* smooth, const float* u_rr,
                  const float* u_mix, const float* bsdf_wi, const float* bsdf_pdf, const float* bsdf_f,
                  uint8_t* rr_killed, uint8_t* guided, uint8_t* mix, float* wi, float* thp) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (!(rr_prob > 0.0f && rr_prob <= 1.0f) || !(bsdf_fraction >= 0.0f && bsdf_fraction <= 1.0f))
        return DFG_ERR_INVALID_ARG;
    if (n == 0) return DFG_OK;
    if (!depth || !smooth || !u_rr || !u_mix || !bsdf_wi || !bsdf_pdf || !bsdf_f || !wi) return DFG_ERR_INVALID_ARG;
    SampleWiArgs a;
    a.rr_prob = rr_prob;
    a.bsdf_fraction = bsdf_fraction;
    a.guiding_enabled = guiding_enabled != 0;
    a.max_guided_depth = max_guided_depth;
    a.depth = depth; a.smooth = smooth; a.u_rr = u_rr; a.u_mix = u_mix;
    a.bsdf_wi = bsdf_wi; a.bsdf_pdf = bsdf_pdf; a.bsdf_f = bsdf_f;
    a.rr_killed = rr_killed; a.guided = guided; a.mix = mix; a.wi = wi; a.thp = thp;
    launch_sample_wi(n, a, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

static int check_producer_records(const dfg_records* r, bool scatter) {
    if (!r) return DFG_ERR_INVALID_ARG;
    if (r->max_depth < 1 || r->n_pixels < 1) return DFG_ERR_BAD_SIZE;
    const void* need[] = {r->cur_depth, r->pos, r->wo, r->normal, r->roughness, r->L, r->thp, r->diffuse,
                          r->specular, r->delta, r->miss};
    for (const void* p : need)
        if (!p) return DFG_ERR_INVALID_ARG;
    if (scatter && (!r->dir || !r->wi_pdf || !r->mis_weight || !r->bsdf_val)) return DFG_ERR_INVALID_ARG;
    return DFG_OK;
}

int dfg_record_scatter(void* stream, int64_t n, dfg_records* rec, const dfg_shading* sd, int train_enabled,
                       int max_train_depth, uint32_t pixel_offset, uint32_t pixel_stride, int mis_aware,
                       int product_with_cosine, const int32_t* pixel, const int32_t* depth, const uint8_t* rr_killed,
                       const uint8_t* alive, const float* wi_dir, const float* thp, const float* scatter_pdf,
                       const float* mis, const float* bsdf_val, const float* cos_theta, const uint8_t* delta) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (pixel_stride < 1) return DFG_ERR_BAD_SIZE;
    if (n == 0) return DFG_OK;
    const int rc = check_producer_records(rec, alive != nullptr);
    if (rc != DFG_OK) return rc;
    if (!sd || !sd->pos || !sd->normal || !sd->wo || !sd->roughness || !sd->diffuse || !sd->specular || !sd->smooth ||
        !pixel || !depth)
        return DFG_ERR_INVALID_ARG;
    if (alive && (!wi_dir || !thp || !scatter_pdf || !bsdf_val || !cos_theta || (mis_aware && !mis)))
        return DFG_ERR_INVALID_ARG;
    RecordScatterArgs a;
    a.train_enabled = train_enabled != 0;
    a.max_train_depth = max_train_depth;
    a.mis_aware = mis_aware != 0;
    a.product_with_cosine = product_with_cosine != 0;
    a.pixel_offset = pixel_offset;
    a.pixel_stride = pixel_stride;
    a.pixel = pixel; a.depth = depth; a.rr_killed = rr_killed; a.alive = alive; a.wi_dir = wi_dir; a.thp = thp;
    a.scatter_pdf = scatter_pdf; a.mis = mis; a.bsdf_val = bsdf_val; a.cos_theta = cos_theta; a.delta = delta;
    launch_record_scatter(n, *rec, *sd, a, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

int dfg_record_radiance(void* stream, int64_t n, dfg_records* rec, int train_enabled, int max_train_depth,
                        uint32_t pixel_offset, uint32_t pixel_stride, const int32_t* pixel, const int32_t* depth,
                        const float* contrib, const float* Le, const uint8_t* miss) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (pixel_stride < 1) return DFG_ERR_BAD_SIZE;
    if (n == 0) return DFG_OK;
    const int rc = check_producer_records(rec, false);
    if (rc != DFG_OK) return rc;
    if (!pixel || !contrib || (miss && !depth)) return DFG_ERR_INVALID_ARG;
    RecordRadianceArgs a;
    a.train_enabled = train_enabled != 0;
    a.max_train_depth = max_train_depth;
    a.pixel_offset = pixel_offset;
    a.pixel_stride = pixel_stride;
    a.pixel = pixel; a.depth = depth; a.contrib = contrib; a.Le = Le; a.miss = miss;
    launch_record_radiance(n, *rec, a, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

}  // extern "C"
