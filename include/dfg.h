/*
 * dfg.h — C ABI of the MI355X-native Distribution-Factorization guiding field.
 *
 * The library owns the three networks of the guide (theta-marginal net,
 * phi|theta-conditional net, radiance cache "NRC"), their fp32 master
 * weights, fp16 training weights, fp16 EMA inference weights and Adam state,
 * all resident in device memory.  Every batch pointer below is a DEVICE
 * pointer owned by the caller; every call is ordered on the hipStream_t passed
 * as `stream` (NULL = default stream) and returns a dfg_status.  No C++
 * exception crosses this boundary.  A context is single-stream serialised:
 * concurrent use needs one context per host thread.
 *
 * Reference surface replaced (pedrovfigueiredo/guiding-df):
 *   LearnedDistribution   src/render/guided/learnedDistribution.h:37-168
 *   NRC                   src/render/guided/nrc.h:30-97
 *   tcnn module API model src/ext/tcnn/include/tiny-cuda-nn/cpp_api.h:54-92
 */
#ifndef DFG_H
#define DFG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfg_ctx dfg_ctx;

typedef enum dfg_status {
    DFG_OK = 0,
    DFG_ERR_INVALID_ARG = 1,   /* null pointer, unknown net id, bad enum        */
    DFG_ERR_BAD_SIZE = 2,      /* n < 0, n above the context maximum, size mismatch */
    DFG_ERR_CONFIG = 3,        /* JSON parse error or unsupported configuration */
    DFG_ERR_HIP = 4,           /* a HIP runtime call failed                     */
    DFG_ERR_OOM = 5,           /* device allocation failed                      */
    DFG_ERR_NOT_INIT = 6,      /* context is NULL / not initialised             */
    DFG_ERR_IO = 7,            /* file open/read/write failed                   */
    DFG_ERR_OVER_BUDGET = 8    /* request exceeds the configured memory budget  */
} dfg_status;

/* Net identifiers for the parameter entry points. */
#define DFG_NET_THETA 0
#define DFG_NET_PHI 1
#define DFG_NET_NRC 2

/* Parameter views for dfg_get_params. */
#define DFG_PARAMS_MASTER 0     /* fp32 master weights (what the optimizer updates) */
#define DFG_PARAMS_TRAIN 1      /* fp16 weights used by the training forward, as fp32 */
#define DFG_PARAMS_INFERENCE 2  /* fp16 EMA weights used by every query, as fp32    */

/* Row layouts (AoS float32) — src/render/guided/guideditem.h:53-110.
 *   DF features  [n][9]  : pos(3) in [0,1], (wo+1)/2 (3), normal theta/pi,
 *                          normal phi/2pi, 1-exp(-roughness)      (GuidedInput)
 *   NRC features [n][15] : the 9 above, diffuse(3), specular(3)  (GuidedInputNRC)
 *   u            [n][4]  : u_theta, u_phi, extra uniform for a degenerate theta
 *                          segment, extra uniform for a degenerate phi segment
 *                          (the reference draws those from sampler.get1D(),
 *                          src/render/guided/train.h:482,494)
 *   wi           [n][4]  : (u0, u1, p_theta, p_phi) — parameters.h:44
 *   target       [B][9]  : dir(2) square coords, L(3), wiPdf, bsdfVal(3) (GuidedOutput)
 */

/* Create a context from a JSON configuration, build the networks and
 * initialise parameters from the configured seed.  Sections: "nn" holds the
 * reference's nn JSON keys (optimizer, encoding, network; "phi_encoding" for
 * the phi net's encoding), "nrc" the NRC net's keys, and "dfg" the guide's own
 * keys: n_bins_theta, n_bins_phi, interpolation ("linear"|"nearest"), seed,
 * n_input_dims, grid_prepass_min_rows.
 * Replaces LearnedDistribution::resetNetwork (learnedDistribution.cpp:64-165)
 * and NRC::resetNetwork (nrc.cpp:35-60). */
int dfg_create(const char* json_config, dfg_ctx** out_ctx);
int dfg_destroy(dfg_ctx* ctx);
/* Last error message of the context (or of the last failed dfg_create when ctx
 * is NULL).  Never NULL. */
const char* dfg_last_error(const dfg_ctx* ctx);

/* Static description of the context. */
typedef struct dfg_info {
    int n_bins_theta, n_bins_phi;   /* bins of the two 1-D PDFs           */
    int interpolation;              /* 1 = linear (AR_LINEAR_INTERPOLATION), 0 = nearest */
    int n_input_dims;               /* DF feature dims (theta net input)  */
    int has_nrc;                    /* 1 if an NRC net is configured      */
    int enc_width[3];               /* padded encoded width per net       */
    int n_hidden[3];                /* hidden layers per net              */
    int n_out_padded[3];            /* padded output width per net        */
    size_t n_params[3];             /* parameters per net (MLP then encoding) */
    size_t n_matrix_params[3];      /* leading MLP (matrix) parameters    */
    int64_t max_batch;              /* largest n accepted by query/train  */
} dfg_info;
int dfg_get_info(const dfg_ctx* ctx, dfg_info* info);

/* Parameters in the reference order: MLP matrices (first, hidden..., last
 * padded to 16 rows, each row-major [out][in]) followed by the encoding's
 * parameters by level offset table
 * (network_with_input_encoding.h:113-155, fully_fused_mlp.cu:873-883,
 *  grid.h:959-1025).  Host pointers, n = dfg_info.n_params[net]. */
int dfg_get_params(dfg_ctx* ctx, int net, int which, float* host_out, size_t n);
/* Gradient of the last training step of `net` (fp32, parameter order), as the
 * optimizer consumed it: loss-scaled by 1024 (TRAIN_LOSS_SCALE, parameters.h:45),
 * i.e. 1024 * d(mean loss)/d(param). */
int dfg_get_grad(dfg_ctx* ctx, int net, float* host_out, size_t n);
/* Sets the fp32 master, the fp16 training copy and the fp16 EMA copy, like
 * Trainer::set_params_full_precision (trainer.h:238-250); resets nothing else. */
int dfg_set_params(dfg_ctx* ctx, int net, const float* host_in, size_t n);
/* Raw little-endian fp32 files <prefix>AR_network_theta.bin,
 * <prefix>AR_network_phi.bin, <prefix>NRC_network.bin — the reference's
 * format (learnedDistribution.cpp:295-382, nrc.cpp:62-112). */
int dfg_save(dfg_ctx* ctx, const char* path_prefix);
int dfg_load(dfg_ctx* ctx, const char* path_prefix);

/* Guided query for one wavefront stage (replaces
 * LearnedDistribution::inferenceStep, learnedDistribution.cpp:167-206, and the
 * thetaPostInference/phiPostInference kernels, train.h:895-929).
 * For each i < n: if guided == NULL or guided[i] != 0, sample (u0, u1) from the
 * guide with u[i]; otherwise keep the direction already in wi[i][0..1] (a BSDF
 * sample).  Then multiply wi[i][2] by p_theta(u0) and wi[i][3] by p_phi(u1|u0)
 * (the reference keeps the sign of its initial +-1 there, train.h:741-750).
 * feat has `feat_stride` floats per row (>= n_input_dims).  When nrc_rgb is
 * non-NULL and an NRC net exists, also writes the NRC radiance [n][3]
 * (feat_stride must then be >= 15).  Fused MFMA path: one query kernel, led
 * for n >= grid_prepass_min_rows by a grid-encoding pre-pass (three kernels). */
int dfg_query(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
              const float* u, const uint8_t* guided, float* wi, float* nrc_rgb);

/* Same contract as dfg_query computed by the plain FP32 per-batch path: one
 * kernel per encoding and per layer, fp32 activations, no MFMA, a host
 * synchronisation per call — the comparator the throughput target is quoted
 * against. */
int dfg_query_plain(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
                    const float* u, const uint8_t* guided, float* wi, float* nrc_rgb);

/* dfg_query without any device allocation during the call.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * contract and bit-identical results as dfg_query, except:
 *   - the grid pre-pass buffers live in one 16.75 MiB device array of the
 *     library (per GPU, held until the process ends), so n is limited to what
 *     fits: 65,536 rows for two 16-level F = 2 grids (or two 8-level F = 4
 *     grids); larger n returns DFG_ERR_BAD_SIZE with the maximum for the
 *     context's configuration;
 *   - the array is shared by every call on a GPU: calls that overlap on one
 *     GPU (two streams, contexts or threads) race; order them on one stream;
 *   - no chunked side-stream pipelining (query_chunks is ignored). */
int dfg_query_noalloc(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
                      const float* u, const uint8_t* guided, float* wi, float* nrc_rgb);

/* Solid-angle pdf of given directions (AutoRegressive::pdf, AR.h:41-43):
 * pdf[i] = p_theta(wi2[i][0]) * p_phi(wi2[i][1] | theta) / (4 pi). */
int dfg_pdf(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
            const float* wi2, float* pdf);

/* One DF training step on B records (LearnedDistribution::trainStep inner
 * batch, learnedDistribution.cpp:230-290): forward of both nets with the
 * training weights, softmax, KL loss and hand-written logits gradient
 * (train.h:931-965, AR.h:50-116), backward, Adam + EMA.  loss_out (device, may
 * be NULL) receives the mean per-record loss. */
int dfg_train_step(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                   const float* target, float* loss_out);

/* NRC inference (NRC::inferenceStep, nrc.cpp:172-181): feat [n][15] -> rgb [n][3]. */
int dfg_nrc_infer(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
                  float* rgb);
/* NRC training step with RelativeL2Luminance (nrc.cpp:114-170):
 * feat [B][15], target_rgb [B][3]; loss_out (device, may be NULL) receives the
 * summed loss divided by B, as the reference reports it.  It has step scratch
 * of its own and touches only the NRC's parameters, so it may be enqueued on
 * another stream beside dfg_train_step (both called from one host thread; the
 * caller orders each stream after the inputs it reads).  The same holds for
 * the two plain steps below. */
int dfg_nrc_train_step(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                       const float* target_rgb, float* loss_out);

/* dfg_train_step and dfg_nrc_train_step computed by the plain FP32 per-batch
 * path: one kernel per encoding and per layer forward and backward, fp32
 * activations, bins and gradients, no MFMA, a host synchronisation per call.
 * They update the same weights, Adam and EMA state.  No counterpart in the
 * reference: they are the comparator the training throughput is quoted
 * against. */
int dfg_train_step_plain(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                         const float* target, float* loss_out);
int dfg_nrc_train_step_plain(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                             const float* target_rgb, float* loss_out);

/* ---- Training records: the step before dfg_train_step ----------------------
 * Path vertices recorded for the training pixels, one SoA slice per depth as
 * GuidedPixelStateBuffer holds them (src/render/guided/workqueue.h:26-144):
 * element (depth d, pixel p) of a field sits at index d * n_pixels + p
 * (vectors as 3 or 2 consecutive floats).  All pointers are device pointers. */
typedef struct dfg_records {
    int max_depth;             /* depth slices (MAX_TRAIN_DEPTH, parameters.h:24)        */
    int64_t n_pixels;          /* pixels per slice                                      */
    uint32_t* cur_depth;       /* [n_pixels] vertices recorded per pixel                */
    float* pos;                /* [3] world-space vertex                                */
    float* dir;                /* [2] spherical (theta, phi) of the scattered direction */
    float* wo;                 /* [2] spherical outgoing direction                      */
    float* normal;             /* [2] spherical shading normal                          */
    float* roughness;          /* [1]                                                   */
    float* L;                  /* [3] radiance gathered after the vertex                */
    float* thp;                /* [3] throughput up to the vertex                       */
    float* bsdf_val;           /* [3]                                                   */
    float* diffuse;            /* [3]                                                   */
    float* specular;           /* [3]                                                   */
    float* wi_pdf;             /* [1] pdf of the scattered direction                    */
    float* mis_weight;         /* [1]                                                   */
    uint8_t* delta;            /* [1] sampled from a delta lobe                         */
    uint8_t* miss;             /* [1] the path left the scene there                     */
    float* Le;                 /* [3] emission seen at the vertex (may be NULL: zero)   */
} dfg_records;
/* Written by: dfg_record_scatter and dfg_record_radiance (the record
 * producers) and dfg_nrc_substitute_radiance (L, cur_depth); read by the
 * row generators and training passes below.  The reference's bsdfPdf and
 * isTwoSided record fields are not kept: nothing on the training path reads
 * them (train.cu:6-213). */

/* Shading data of the scatter queue's rows (the ShadingData fields the
 * records keep, shared.h:36-52), device arrays [n][k]. */
typedef struct dfg_shading {
    const float* pos;          /* [3] hit point                                         */
    const float* normal;       /* [3] shading normal (frame.N)                          */
    const float* wo;           /* [3] outgoing direction                                */
    const float* roughness;    /* [1]                                                   */
    const float* diffuse;      /* [3]                                                   */
    const float* specular;     /* [3]                                                   */
    const uint8_t* smooth;     /* [1] getBsdfType() has a diffuse or glossy lobe       */
} dfg_shading;

/* ---- Producing the records in a wavefront renderer --------------------------
 * Training pixels are those with (pixel - pixel_offset) % pixel_stride == 0 in
 * uint32 arithmetic (isTrainingPixel, guided.h:26-28); train_enabled stands
 * for the reference's enableTraining || oneStep.  Queue rows whose pixel lies
 * outside [0, rec->n_pixels) are ignored.  A pixel may appear at most once
 * per call (one ray per pixel and bounce); calls on one stream apply in order.
 *
 * The record step of GuidedPathTracer::handleIntersections
 * (integrator.cpp:277-296) and handleNRCLastBounce (integrator.cpp:193-201),
 * after dfg_scatter.  For each queue row i of a training pixel with
 * depth[i] < max_train_depth (and rr_killed[i] == 0 when given):
 *   alive[i] != 0 -> incrementDepth (workqueue.h:26-58): the vertex at
 *     cur_depth gets L = 0, thp = thp[i] (the updated throughput), pos,
 *     dir = spherical(wi_dir[i]), wo, normal (spherical, math_utils.h:183-190),
 *     roughness, diffuse, specular, wiPdf = scatter_pdf[i], misWeight =
 *     mis_aware ? mis[i] : 1, delta[i], miss = 0, Le = 0, bsdfVal =
 *     product_with_cosine ? bsdf_val |cos| : bsdf_val; cur_depth + 1;
 *   otherwise (alive NULL: every row) -> incrementDepthLastBounceNRC
 *     (workqueue.h:60-81): pos, normal, wo, roughness, diffuse, specular,
 *     delta = !smooth, miss = 0, L = thp = Le = 0; cur_depth + 1.
 * Nothing is written at cur_depth >= rec->max_depth.  rr_killed and delta may
 * be NULL (all 0). */
int dfg_record_scatter(void* stream, int64_t n, dfg_records* rec, const dfg_shading* sd, int train_enabled,
                       int max_train_depth, uint32_t pixel_offset, uint32_t pixel_stride, int mis_aware,
                       int product_with_cosine, const int32_t* pixel, const int32_t* depth, const uint8_t* rr_killed,
                       const uint8_t* alive, const float* wi_dir, const float* thp, const float* scatter_pdf,
                       const float* mis, const float* bsdf_val, const float* cos_theta, const uint8_t* delta);
/* Radiance reaching the path (handleEmissiveHit / handleMiss,
 * integrator.cpp:144-191).  For each row i of a training pixel:
 * recordRadiance (workqueue.h:83-92) adds contrib[i] (the radiance times the
 * path throughput) to L of every recorded vertex below min(cur_depth,
 * max_depth); overwriteLe (workqueue.h:98-106) sets Le of the last vertex to
 * Le[i] (Le may be NULL: skipped); and where miss[i] != 0 and depth[i] <=
 * max_train_depth, a placeholder vertex (incrementDepthLastBounceNRC with an
 * empty ShadingData: miss = 1, zero position and reflectances, roughness 1)
 * keeps the radiance-cache inputs aligned.  miss may be NULL. */
int dfg_record_radiance(void* stream, int64_t n, dfg_records* rec, int train_enabled, int max_train_depth,
                        uint32_t pixel_offset, uint32_t pixel_stride, const int32_t* pixel, const int32_t* depth,
                        const float* contrib, const float* Le, const uint8_t* miss);

/* DF training rows (generate_training_data_AR, train.cu:6-96).  For the
 * training pixels p = pixel_offset + t * pixel_stride, t < n_train, every
 * recorded vertex that is not delta, has finite inputs and a non-zero wiPdf
 * becomes one row: feat [9] (as dfg_prepare_features builds them) and target
 * [9] = (wi square coords (2), L / thp * mis * bsdfVal (3), wiPdf, bsdfVal
 * slot left 0 as in the reference).  Rows are written pixel-major, depth
 * inner (a deterministic order; the reference appends through an atomic).
 * aabb[6] = scene (min xyz, max xyz), device.  *count (device int64) gets the
 * number of valid rows; rows past `capacity` are not written. */
int dfg_training_targets(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train,
                         uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb, float* feat,
                         float* target, int64_t capacity, int64_t* count);
/* NRC training rows (generate_training_data_nrc, train.cu:98-151): vertices
 * before the last of each pixel, neither delta nor miss: feat [15] (the 9 DF
 * features, diffuse, specular), target_rgb [3] = L / thp * mis * bsdfVal /
 * (wiPdf + 1e-5). */
int dfg_nrc_training_targets(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train,
                             uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb, float* feat,
                             float* target_rgb, int64_t capacity, int64_t* count);
/* One DF training pass from records (LearnedDistribution::trainStep,
 * learnedDistribution.cpp:208-293): rows into the context's training buffer,
 * then min(rows / batch_size + 1, batch_per_frame) steps on consecutive
 * slices of batch_size rows, each rounded down to a multiple of 128; a slice
 * under 65,536 rows ends the pass.  Synchronises `stream`; on the host:
 * n_rows, n_steps run, and the reference's loss (sum of the per-step mean
 * losses divided by the planned step count). */
int dfg_train_records(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train,
                      uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb, int64_t batch_size,
                      int batch_per_frame, int64_t* n_rows, int* n_steps, float* loss);
/* NRC::trainStep (nrc.cpp:114-170) from records: its own schedule, max(rows /
 * batch_size + 1, batch_per_frame) steps of min(rows / steps, batch_size)
 * rows rounded down to 128, slice k starting at k times its size. */
int dfg_nrc_train_records(dfg_ctx* ctx, void* stream, const dfg_records* rec, int64_t n_train,
                          uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb, int64_t batch_size,
                          int batch_per_frame, int64_t* n_rows, int* n_steps, float* loss);

/* Radiance-cache substitution of the training records (the reference's
 * OPTIMIZE_ON_NRC = 4 path: generate_training_data_nrc_with_inference,
 * train.cu:156-213, NRC::inferenceStepFromTrainingData, nrc.cpp:183-216,
 * apply_NRCoutput_to_training_data and divide_training_data_by_NRCoutput,
 * train.cu:265-308).  For every recorded vertex k < cur_depth that is neither
 * delta nor miss, the NRC radiance R_k is inferred; then cur_depth drops by one
 * per pixel; then, for each such k, the first non-delta vertex j <= k - 1 gets
 * L_j = thp_{k-1} * (R_k + Le_{k-1}); then L_k /= R_k per channel where
 * R_k > 1e-5 (only for vertices still below cur_depth).  Writes rec->L and
 * rec->cur_depth.  Run it after the NRC training pass and before the DF one,
 * as the reference's trainStep does (integrator.cpp:1202-1237).  Synchronises
 * `stream`; *n_inferred (host, may be NULL) gets the number of NRC queries. */
int dfg_nrc_substitute_radiance(dfg_ctx* ctx, void* stream, dfg_records* rec, int64_t n_train,
                                uint32_t pixel_offset, uint32_t pixel_stride, const float* aabb,
                                int64_t* n_inferred);

/* Dense pdf image of one shading point (LearnedDistribution::
 * mouseOverInference / mouseOverInferenceOptimized, learnedDistribution.cpp:
 * 386-587): image[i][j] = the solid-angle pdf (as dfg_pdf) at the square
 * coordinates u0 = (i + 0.5) / res_theta, u1 = (j + 0.5) / res_phi.  feat:
 * device [n_input_dims].  res_theta, res_phi in 1..4096. */
int dfg_pdf_image(dfg_ctx* ctx, void* stream, const float* feat, int res_theta, int res_phi, float* image);

/* ---- Around the query in a wavefront renderer -------------------------------
 * Directions and solid-angle pdfs of query results (AutoRegressive::sample and
 * pdf, AR.h:25-48): dir[i] = squareToUniformSphere(wi[i][0], wi[i][1])
 * (math_utils.h:226-231); pdf[i] = |p_theta| |p_phi| / (4 pi).  dir or pdf may
 * be NULL. */
int dfg_directions(void* stream, int64_t n, const float* wi, float* dir, float* pdf);
/* The decision before the query (GuidedPathTracer::sampleWi,
 * integrator.cpp:355-401), with the caller's uniforms and BSDF sample:
 *   rr_killed[i] = u_rr[i] >= rr_prob (Russian roulette; such a row is left
 *     as it is, guided = mix = 0);
 *   otherwise thp[i] /= rr_prob (the w.thp /= probRR of handleIntersections,
 *     integrator.cpp:211-212; thp may be NULL), and with on = guiding_enabled
 *     && depth[i] < max_guided_depth && smooth[i]:
 *     guided[i] = on && (bsdf_fraction == 0 || u_mix[i] >= bsdf_fraction);
 *     mix[i] = !guided[i] && on && bsdf_fraction < 1 (the BSDF-sampled
 *       direction takes the guide's pdf into the mixture, integrator.cpp:256-261);
 *     wi[i][2..3] = 1, and for !guided rows whose BSDF sample has a non-zero
 *       pdf and value, wi[i][0..1] = uniformSphereToSquare(bsdf_wi[i])
 *       (math_utils.h:233-240), the direction dfg_query then evaluates.
 * The reference draws u_mix only when bsdf_fraction != 0; here the caller
 * supplies both uniforms.  bsdf_wi [n][3] world direction, bsdf_f [n][3]. */
int dfg_sample_wi(void* stream, int64_t n, float rr_prob, float bsdf_fraction, int guiding_enabled,
                  int max_guided_depth, const int32_t* depth, const uint8_t* smooth, const float* u_rr,
                  const float* u_mix, const float* bsdf_wi, const float* bsdf_pdf, const float* bsdf_f,
                  uint8_t* rr_killed, uint8_t* guided, uint8_t* mix, float* wi, float* thp);
/* The scatter step of GuidedPathTracer::handleIntersections
 * (integrator.cpp:203-299) once the BSDF is evaluated: per row, the guide pdf
 * pg = |p_theta| |p_phi| / (4 pi) from wi [n][4], the BSDF pdf pb, value
 * bsdf_val [n][3] and |cos| at the scattered direction come from the caller.
 * With f = bsdf_fraction: guided rows scatter with (1 - f) pg + f pb; other
 * rows with pb, or with f pb + (1 - f) pg when mix[i] != 0 (guiding on, f < 1
 * and a smooth BSDF).  Where that pdf is non-zero and bsdf_val is not all
 * zero: thp[i] *= bsdf_val |cos| / (pdf + 1e-5), alive[i] = 1, and
 * mis[i] = (1 - f) pg / pdf (the training records' MIS weight); otherwise
 * alive[i] = 0 and thp is left.  scatter_pdf [n] receives the pdf.  guided and
 * mix may be NULL (all 0). */
int dfg_scatter(void* stream, int64_t n, float bsdf_fraction, const float* wi, const uint8_t* guided,
                const uint8_t* mix, const float* bsdf_pdf, const float* bsdf_val, const float* cos_theta,
                float* thp, float* scatter_pdf, float* mis, uint8_t* alive);

/* Feature preparation (generate_inference_data_AR, train.cu:334-362):
 * shading points -> DF feature rows [n][9].
 *   pos [n][3] world space, aabb[6] = (min xyz, max xyz), wo [n][3] (any length),
 *   normal [n][3] (any length), roughness [n]. */
int dfg_prepare_features(void* stream, int64_t n, const float* pos, const float* aabb,
                         const float* wo, const float* normal, const float* roughness,
                         float* feat);

/* Stateless DF head on given logits (softmax x N, inverse-CDF sample, pdf):
 * the thetaPostInference / phiPostInference math (train.h:844-929) without the
 * networks.  logits_theta [n][n_theta], logits_phi [n][n_phi]; interpolation
 * 1 = linear, 0 = nearest.  Outputs (each may be NULL): bins [n][N] (softmax x N),
 * seg [n][2] (sampled CDF segment of theta and phi; -1 when not sampled).
 * The phi logits are taken as given (not recomputed from the sampled theta). */
int dfg_df_head(void* stream, int64_t n, int n_theta, int n_phi, int interpolation,
                const float* logits_theta, const float* logits_phi, const float* u,
                const uint8_t* guided, float* wi, float* bins_theta, float* bins_phi,
                int32_t* seg);

/* ---- Batched wavefront (all bounces of a frame in one persistent launch) ----
 * Replaces the per-bounce LearnedDistribution::inferenceStep calls of the
 * bounce loop (integrator.cpp:588-633, one call + device sync per bounce) for
 * a frame whose bounce batches are packed back to back: bounce b occupies rows
 * [sum_{k<b} sizes[k], +sizes[b]) of feat/u/guided/wi.  The sizes are read on
 * the device, so the captured hipGraph replays for any sizes whose sum stays
 * within the plan's capacity. */
typedef struct dfg_wavefront dfg_wavefront;

typedef struct dfg_wavefront_tuning {
    int blocks_per_cu;     /* resident wavefronts launched per CU (persistent grid) */
    int tiles_per_fetch;   /* 64-row tiles taken per work-counter fetch            */
    int grid_prepass;      /* 1: XCD-aware grid pre-pass nodes before the kernel   */
    float ms;              /* autotune: measured frame time of this choice         */
} dfg_wavefront_tuning;

int dfg_wavefront_create(dfg_ctx* ctx, int max_bounces, int64_t capacity, dfg_wavefront** out);
int dfg_wavefront_destroy(dfg_wavefront* w);
int dfg_wavefront_set_tuning(dfg_wavefront* w, const dfg_wavefront_tuning* t);
int dfg_wavefront_get_tuning(const dfg_wavefront* w, dfg_wavefront_tuning* t);
/* sizes: DEVICE int64 [n_bounces].  Captures the frame graph on first use (and
 * again when a buffer pointer, the bounce count or the tuning changes), then
 * launches it on `stream`. */
int dfg_wavefront_run(dfg_wavefront* w, void* stream, int n_bounces, const int64_t* sizes, const float* feat,
                      int feat_stride, const float* u, const uint8_t* guided, float* wi);
/* Times every candidate (persistent grid size x tiles per fetch x pre-pass) on
 * the given frame and keeps the fastest; writes it to *best (may be NULL). */
int dfg_wavefront_autotune(dfg_wavefront* w, void* stream, int n_bounces, const int64_t* sizes, const float* feat,
                           int feat_stride, const float* u, const uint8_t* guided, float* wi,
                           dfg_wavefront_tuning* best);

/* ---- Served loop: continuous batching of guidance queries ----------------
 * dfg_server_submit puts a request (device-resident rows) into a request ring
 * in pinned host memory.  dfg_server_progress launches rounds, each one
 * captured graph on the next free slot (two per stream, `n_streams` streams),
 * and retires finished ones.  Admission runs on the device: the first kernel
 * of a round takes requests from the ring's head, in arrival order, while the
 * rows in flight (a device counter that every round's last kernel gives back)
 * stay within the memory budget and the round within `round_rows`, and builds
 * the round's work list there; the round then runs the grid pre-pass and one
 * persistent work-list query over every admitted row.  A round started while
 * another runs overlaps it on the device and admits what arrived meanwhile.
 * A queued training step is interleaved after every `train_every` query
 * batches (requests; enqueued behind the round that completes the count,
 * ordered after the rounds before it and before the ones after it).  Each
 * request is timed from submission to the completion event of its round.
 * Each row's result is identical whether its request ran alone or in a mix.
 * No counterpart in the reference (its queries are one batch per bounce). */
typedef struct dfg_server dfg_server;

typedef struct dfg_server_config {
    int64_t budget_bytes;   /* device bytes of request rows admitted at once        */
    int max_batch_requests; /* requests packed into one launch (<= 256)             */
    int n_streams;          /* launch streams (1..4)                                */
    int train_every;        /* query batches between training steps (0: never)     */
    int64_t round_rows;     /* rows of one round, the largest request accepted; also
                               the per-slot pre-pass capacity (0: max(2^20,
                               min(budget rows, 2^22)))                              */
} dfg_server_config;

typedef struct dfg_request {
    int64_t id;
    int64_t n;
    const float* feat;      /* [n][feat_stride] */
    int feat_stride;
    const float* u;         /* [n][4] */
    const uint8_t* guided;  /* [n] or NULL */
    float* wi;              /* [n][4] in/out */
    float* nrc_rgb;         /* [n][3] or NULL (all requests of a server agree) */
} dfg_request;

typedef struct dfg_server_stats {
    int64_t submitted, completed, launches, train_steps, rows;
    double p50_ms, p99_ms, mean_ms, max_ms;  /* request latency, submit -> done */
    double busy_s;                            /* first submit -> last completion */
    int64_t peak_rows;                        /* most rows in flight after an admission */
} dfg_server_stats;

int dfg_server_create(dfg_ctx* ctx, const dfg_server_config* cfg, dfg_server** out);
int dfg_server_destroy(dfg_server* s);
int dfg_server_submit(dfg_server* s, const dfg_request* r);
/* Training step queued for the next interleave point: DF records (and, if
 * nrc_feat is non-NULL, an NRC step on nrc_feat/nrc_target). */
int dfg_server_submit_train(dfg_server* s, int64_t B, const float* feat, int feat_stride, const float* target,
                            const float* nrc_feat, const float* nrc_target);
/* wait = 0: admit/launch/retire what is ready and return; 1: until idle. */
int dfg_server_progress(dfg_server* s, int wait);
int dfg_server_get_stats(dfg_server* s, dfg_server_stats* out);
/* Per-request records of completed requests (drained): id, latency (ms) and
 * the number of training steps applied before the request's launch. */
int dfg_server_drain(dfg_server* s, int64_t* ids, double* latency_ms, int64_t* weights_version, int max,
                     int* count);

/* Debug exports.  dfg_logits: raw logits of both nets (inference weights) for
 * rows feat[n] with the phi net fed theta_in[n] (square u0 of theta).
 * dfg_encode: encoded (padded) input of net `net` [n][enc_width] for raw inputs
 * in [n][in_dims] (phi: n_input_dims + 1, the last one the theta input).
 * dfg_grid_index: hash/dense grid table indices [n][levels][8] of net `net`. */
int dfg_logits(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride,
               const float* theta_in, float* logits_theta, float* logits_phi);
int dfg_encode(dfg_ctx* ctx, void* stream, int net, int64_t n, const float* in, int in_stride,
               float* out);
int dfg_grid_index(dfg_ctx* ctx, void* stream, int net, int64_t n, const float* in,
                   int in_stride, uint32_t* idx);

#ifdef __cplusplus
}
#endif
#endif /* DFG_H */
