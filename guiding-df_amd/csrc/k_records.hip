// Training rows from recorded path vertices: the step before the training
// step (generate_training_data_AR, train.cu:6-96, and
// generate_training_data_nrc, train.cu:98-151).  Rows are compacted in a
// deterministic order (training pixel t, then depth) by count / scan / write:
//   count : per 256-pixel block, valid vertices per pixel and their in-block
//           exclusive offsets; the block total
//   scan  : one block, exclusive offsets of the block totals; the row count
//   write : each pixel writes its rows at block offset + in-block offset
#include "../../include/dfg.h"
#include "device.h"
#include "internal.h"

namespace dfg {

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647693f;
constexpr float kEps = 1e-5f;  // M_EPSILON, constants.h:19

// sphericalToCartesian, math_utils.h:198-208
__device__ __forceinline__ void sph_to_cart(float theta, float phi, float (&v)[3]) {
    const float st = sinf(theta), ct = cosf(theta), sp = sinf(phi), cp = cosf(phi);
    v[0] = st * cp;
    v[1] = st * sp;
    v[2] = ct;
}

struct Row {
    float feat[15];
    float tgt[9];
};

// One record (depth k, pixel p) as a training row; false when the reference
// skips it.  mode 0: DF row (feat 9, target 9); 1: NRC row (feat 15, target 3);
// 2: NRC inference input (feat 15; the NRC row's validity).
__device__ bool record_row(const dfg_records& r, int k, int64_t p, const float* __restrict__ aabb, int mode, Row& o) {
    const int64_t i = (int64_t)k * r.n_pixels + p;
    if (r.delta[i]) return false;
    if (mode != 0 && r.miss[i]) return false;
    // normalizeSpatialCoord (train.h:272-276): the box inflated by 0.5 % of its diagonal
    float d[3];
    for (int c = 0; c < 3; ++c) d[c] = aabb[3 + c] - aabb[c];
    const float infl = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * 0.005f;
    for (int c = 0; c < 3; ++c) {
        const float lo = aabb[c] - infl, hi = aabb[3 + c] + infl;
        o.feat[c] = 0.5f + (r.pos[3 * i + c] - (lo + hi) * 0.5f) / (hi - lo);
    }
    // warp_direction_for_sh of the recorded wo (train.h:1243-1245)
    float w[3];
    sph_to_cart(r.wo[2 * i], r.wo[2 * i + 1], w);
    for (int c = 0; c < 3; ++c) o.feat[3 + c] = (w[c] + 1.0f) * 0.5f;
    o.feat[6] = r.normal[2 * i] * (1.0f / kPi);
    o.feat[7] = r.normal[2 * i + 1] * (1.0f / kTwoPi);
    o.feat[8] = 1.0f - expf(-r.roughness[i]);  // warp_roughness_for_ob, train.h:1247-1249
    bool bad = false;
    for (int c = 0; c < 6; ++c) bad |= o.feat[c] != o.feat[c];
    // local radiance: L / thp per channel where thp > eps, times the MIS weight
    float L[3];
    const float mis = r.mis_weight[i];
    for (int c = 0; c < 3; ++c) {
        const float t = r.thp[3 * i + c];
        L[c] = (t > kEps ? r.L[3 * i + c] / t : 0.0f) * mis;
    }
    const float pdf = r.wi_pdf[i];
    if (mode == 0) {
        // uniformSphereToSquare(sphericalToCartesian(dir)), math_utils.h:233-240
        float v[3];
        sph_to_cart(r.dir[2 * i], r.dir[2 * i + 1], v);
        const float n = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; ++c) v[c] /= n;
        float ph = atan2f(v[1], v[0]) * (1.0f / kTwoPi);
        if (ph < 0.0f) ph += 1.0f;
        o.tgt[0] = (1.0f - v[2]) * 0.5f;
        o.tgt[1] = ph;
        // the reference tests L before the bsdfVal product (train.cu:80-84)
        bad |= o.tgt[0] != o.tgt[0] || o.tgt[1] != o.tgt[1] || pdf != pdf || pdf == 0.0f;
        for (int c = 0; c < 3; ++c) bad |= L[c] != L[c];
        for (int c = 0; c < 3; ++c) o.tgt[2 + c] = L[c] * r.bsdf_val[3 * i + c];
        o.tgt[5] = pdf;
        o.tgt[6] = o.tgt[7] = o.tgt[8] = 0.0f;
    } else {
        for (int c = 0; c < 3; ++c) {
            o.feat[9 + c] = r.diffuse[3 * i + c];
            o.feat[12 + c] = r.specular[3 * i + c];
            o.tgt[c] = L[c] * r.bsdf_val[3 * i + c] / (pdf + kEps);
            bad |= o.tgt[c] != o.tgt[c];
        }
    }
    return !bad;
}

__device__ __forceinline__ int pixel_depth(const dfg_records& r, int64_t p, int mode) {
    int d = (int)r.cur_depth[p];
    if (d > r.max_depth) d = r.max_depth;
    // the NRC training rows have no radiance estimate for the last recorded vertex
    return mode == 1 ? d - 1 : d;
}

__global__ __launch_bounds__(256) void records_count_kernel(const dfg_records r, int64_t n_train, uint32_t offset,
                                                            uint32_t stride, const float* __restrict__ aabb, int mode,
                                                            uint32_t* __restrict__ local, int64_t* __restrict__ block_tot) {
    __shared__ uint32_t s[256];
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t cnt = 0;
    if (t < n_train) {
        const int64_t p = (int64_t)offset + t * (int64_t)stride;
        const int depth = pixel_depth(r, p, mode);
        Row o;
        for (int k = 0; k < depth; ++k) cnt += record_row(r, k, p, aabb, mode, o) ? 1u : 0u;
    }
    s[threadIdx.x] = cnt;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const uint32_t y = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    if (t < n_train) local[t] = s[threadIdx.x] - cnt;
    if (threadIdx.x == 255) block_tot[blockIdx.x] = s[255];
}

// one block: block totals -> exclusive offsets; the row count to *count
__global__ __launch_bounds__(1024) void records_scan_kernel(int64_t* __restrict__ v, int64_t m,
                                                            int64_t* __restrict__ count) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    int64_t b = (int64_t)t * per;
    if (b > m) b = m;
    int64_t e = b + per;
    if (e > m) e = m;
    int64_t sum = 0;
    for (int64_t i = b; i < e; ++i) sum += v[i];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int64_t y = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    int64_t run = part[t] - sum;
    for (int64_t i = b; i < e; ++i) {
        const int64_t c = v[i];
        v[i] = run;
        run += c;
    }
    if (t == 1023) *count = part[1023];
}

// mode 2 writes the vertex (depth, pixel) of each row to vidx instead of a
// target, then lowers the pixel's cur_depth by one (read before, by this
// thread alone)
__global__ __launch_bounds__(256) void records_write_kernel(const dfg_records r, int64_t n_train, uint32_t offset,
                                                            uint32_t stride, const float* __restrict__ aabb, int mode,
                                                            const uint32_t* __restrict__ local,
                                                            const int64_t* __restrict__ block_off,
                                                            float* __restrict__ feat, float* __restrict__ tgt,
                                                            int2* __restrict__ vidx, int64_t capacity) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n_train) return;
    const int64_t p = (int64_t)offset + t * (int64_t)stride;
    const int depth = pixel_depth(r, p, mode);
    int64_t at = block_off[blockIdx.x] + local[t];
    const int nf = mode == 0 ? 9 : 15, nt = mode == 0 ? 9 : 3;
    Row o;
    for (int k = 0; k < depth; ++k) {
        if (!record_row(r, k, p, aabb, mode, o)) continue;
        if (at < capacity) {
            for (int c = 0; c < nf; ++c) feat[at * nf + c] = o.feat[c];
            if (mode == 2) vidx[at] = make_int2(k, (int)p);
            else
                for (int c = 0; c < nt; ++c) tgt[at * nt + c] = o.tgt[c];
        }
        ++at;
    }
    // the reference's unsigned decrement would wrap a pixel with no vertex to
    // 2^32 - 1; it stays at 0 here
    if (mode == 2 && r.cur_depth[p] > 0u) r.cur_depth[p] = r.cur_depth[p] - 1u;
}

// GuidedPixelStateBuffer::isValidRecord (workqueue.h:113-115)
__device__ __forceinline__ bool valid_record(const dfg_records& r, int64_t p, int d) {
    return d >= 0 && d < r.max_depth && (uint32_t)d < r.cur_depth[p];
}

// replaceIndirectRadianceSkipDelta (workqueue.h:131-138) for each inferred
// vertex k: the first non-delta vertex at or below k - 1 takes
// thp_{k-1} * (R_k + Le_{k-1})
__global__ void nrc_apply_kernel(const dfg_records r, const int2* __restrict__ vidx, const float* __restrict__ rgb,
                                 const int64_t* __restrict__ count, int64_t capacity) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t m = *count < capacity ? *count : capacity;
    if (i >= m) return;
    const int k = vidx[i].x;
    const int64_t p = vidx[i].y;
    const int init = k - 1;
    int d = init;
    while (d >= 0 && r.delta[(int64_t)d * r.n_pixels + p]) --d;
    if (!valid_record(r, p, d)) return;
    const int64_t a = (int64_t)init * r.n_pixels + p, b = (int64_t)d * r.n_pixels + p;
    for (int c = 0; c < 3; ++c) {
        const float le = r.Le ? r.Le[3 * a + c] : 0.0f;
        r.L[3 * b + c] = r.thp[3 * a + c] * (rgb[3 * i + c] + le);
    }
}

// divideRadiance (workqueue.h:117-123): L_k /= R_k where R_k > eps
__global__ void nrc_divide_kernel(const dfg_records r, const int2* __restrict__ vidx, const float* __restrict__ rgb,
                                  const int64_t* __restrict__ count, int64_t capacity) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t m = *count < capacity ? *count : capacity;
    if (i >= m) return;
    const int k = vidx[i].x;
    const int64_t p = vidx[i].y;
    if (!valid_record(r, p, k)) return;
    const int64_t a = (int64_t)k * r.n_pixels + p;
    for (int c = 0; c < 3; ++c) {
        const float R = rgb[3 * i + c];
        if (R > kEps) r.L[3 * a + c] = r.L[3 * a + c] / R;
    }
}

}  // namespace

namespace {

// AutoRegressive::sample / pdf (AR.h:25-48) with squareToUniformSphere
// (math_utils.h:226-231)
__global__ void directions_kernel(int64_t n, const float* __restrict__ wi, float* __restrict__ dir,
                                  float* __restrict__ pdf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 w = *(const float4*)(wi + 4 * i);
    if (dir) {
        const float z = 1.0f - 2.0f * w.x;
        const float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
        const float ph = kTwoPi * w.y;
        dir[3 * i + 0] = r * cosf(ph);
        dir[3 * i + 1] = r * sinf(ph);
        dir[3 * i + 2] = z;
    }
    if (pdf) pdf[i] = fabsf(w.z) * fabsf(w.w) * DFG_INV_4PI;
}

// the mixture pdf and throughput update of handleIntersections (integrator.cpp:203-299)
__global__ void scatter_kernel(int64_t n, float f, const float* __restrict__ wi, const uint8_t* __restrict__ guided,
                               const uint8_t* __restrict__ mix, const float* __restrict__ bsdf_pdf,
                               const float* __restrict__ bsdf_val, const float* __restrict__ cos_theta,
                               float* __restrict__ thp, float* __restrict__ spdf, float* __restrict__ mis,
                               uint8_t* __restrict__ alive) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float pg = fabsf(wi[4 * i + 2]) * fabsf(wi[4 * i + 3]) * DFG_INV_4PI;
    const float pb = bsdf_pdf[i];
    float p;
    if (guided && guided[i]) p = (1.0f - f) * pg + f * pb;
    else if (mix && mix[i]) p = pb * f + (1.0f - f) * pg;
    else p = pb;
    const float b0 = bsdf_val[3 * i], b1 = bsdf_val[3 * i + 1], b2 = bsdf_val[3 * i + 2];
    const bool ok = p != 0.0f && (b0 != 0.0f || b1 != 0.0f || b2 != 0.0f);
    if (spdf) spdf[i] = p;
    if (alive) alive[i] = ok ? 1 : 0;
    if (!ok) return;
    const float c = fabsf(cos_theta[i]);
    thp[3 * i + 0] = thp[3 * i + 0] * b0 * c / (p + kEps);
    thp[3 * i + 1] = thp[3 * i + 1] * b1 * c / (p + kEps);
    thp[3 * i + 2] = thp[3 * i + 2] * b2 * c / (p + kEps);
    if (mis) mis[i] = (1.0f - f) * pg / p;
}

// rows of a pdf image: the point's features broadcast, (u0, u1) at pixel centres
__global__ void pdf_image_rows_kernel(const float* __restrict__ feat, int D, int rt, int rp, float* __restrict__ rows,
                                      float* __restrict__ wi2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rt * rp) return;
    for (int d = 0; d < D; ++d) rows[i * D + d] = feat[d];
    const int a = (int)(i / rp), b = (int)(i % rp);
    wi2[2 * i] = ((float)a + 0.5f) / (float)rt;
    wi2[2 * i + 1] = ((float)b + 0.5f) / (float)rp;
}

// ---------------------------------------------------------------------------
// The consumer stage around the query (GuidedPathTracer, integrator.cpp):
// the Russian-roulette and guide-or-BSDF decision of sampleWi, and the
// training-record producers of GuidedPixelStateBuffer (workqueue.h:26-92).
// One thread per queue row; rows whose pixel is outside the records are
// ignored; a pixel appears at most once per launch (one ray per pixel and
// bounce), so the per-pixel updates do not race.
// ---------------------------------------------------------------------------

// v.normalized() as Eigen computes it: v / |v|, v itself when |v| = 0
__device__ __forceinline__ void normalized3(const float* __restrict__ v, float (&o)[3]) {
    const float n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float n = sqrtf(n2);
    for (int c = 0; c < 3; ++c) o[c] = n2 > 0.0f ? v[c] / n : v[c];
}

// cartesianToSpherical, math_utils.h:183-190: (acos z, atan2(y, x) in [0, 2pi))
__device__ __forceinline__ void cart_to_sph(const float* __restrict__ v, float* __restrict__ out) {
    float vn[3];
    normalized3(v, vn);
    const float th = acosf(fminf(fmaxf(vn[2], -1.0f), 1.0f));  // safe_acos
    float ph = atan2f(vn[1], vn[0]);
    if (ph < 0.0f) ph += kTwoPi;
    out[0] = th;
    out[1] = ph;
}

// uniformSphereToSquare, math_utils.h:233-240
__device__ __forceinline__ void sphere_to_square(const float* __restrict__ v, float& u0, float& u1) {
    float vn[3];
    normalized3(v, vn);
    const float ph = atan2f(vn[1], vn[0]) * (1.0f / kTwoPi);
    u0 = (1.0f - vn[2]) * 0.5f;
    u1 = ph < 0.0f ? ph + 1.0f : ph;
}

// isTrainingPixel, guided.h:26-28 (uint32 arithmetic as the reference's)
__device__ __forceinline__ bool training_pixel(uint32_t p, uint32_t offset, uint32_t stride) {
    return (p - offset) % stride == 0u;
}

// sampleWi (integrator.cpp:355-401) and the w.thp /= probRR of
// handleIntersections (integrator.cpp:211-212)
__global__ void sample_wi_kernel(int64_t n, SampleWiArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool killed = a.u_rr[i] >= a.rr_prob;
    if (a.rr_killed) a.rr_killed[i] = killed ? 1 : 0;
    if (killed) {
        if (a.guided) a.guided[i] = 0;
        if (a.mix) a.mix[i] = 0;
        return;
    }
    if (a.thp)
        for (int c = 0; c < 3; ++c) a.thp[3 * i + c] = a.thp[3 * i + c] / a.rr_prob;
    const bool on = a.guiding_enabled && a.depth[i] < a.max_guided_depth && a.smooth[i] != 0;
    const bool guided = on && (a.bsdf_fraction == 0.0f || a.u_mix[i] >= a.bsdf_fraction);
    if (a.guided) a.guided[i] = guided ? 1 : 0;
    // a BSDF-sampled direction still gets the guide's pdf for the mixture
    // when guiding is on at this depth, f < 1 and the BSDF is smooth
    // (integrator.cpp:256-261)
    if (a.mix) a.mix[i] = (!guided && on && a.bsdf_fraction < 1.0f) ? 1 : 0;
    float* w = a.wi + 4 * i;
    w[2] = 1.0f;
    w[3] = 1.0f;
    if (!guided) {
        const float* bf = a.bsdf_f + 3 * i;
        if (a.bsdf_pdf[i] != 0.0f && (bf[0] != 0.0f || bf[1] != 0.0f || bf[2] != 0.0f))
            sphere_to_square(a.bsdf_wi + 3 * i, w[0], w[1]);
    }
}

// incrementDepthLastBounceNRC (workqueue.h:60-81) of shading row i (or of an
// empty ShadingData when sd_row < 0: zero position, reflectances and
// directions, roughness 1, a diffuse BSDF)
__device__ void last_bounce_record(const dfg_records& r, const dfg_shading& sd, int64_t sd_row, uint32_t p,
                                   bool miss) {
    const uint32_t d = r.cur_depth[p];
    if (d >= (uint32_t)r.max_depth) return;
    const int64_t k = (int64_t)d * r.n_pixels + p;
    r.cur_depth[p] = d + 1;
    if (sd_row >= 0) {
        for (int c = 0; c < 3; ++c) {
            r.pos[3 * k + c] = sd.pos[3 * sd_row + c];
            r.diffuse[3 * k + c] = sd.diffuse[3 * sd_row + c];
            r.specular[3 * k + c] = sd.specular[3 * sd_row + c];
        }
        cart_to_sph(sd.normal + 3 * sd_row, r.normal + 2 * k);
        cart_to_sph(sd.wo + 3 * sd_row, r.wo + 2 * k);
        r.roughness[k] = sd.roughness[sd_row];
        r.delta[k] = sd.smooth[sd_row] ? 0 : 1;
    } else {
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        for (int c = 0; c < 3; ++c) r.pos[3 * k + c] = r.diffuse[3 * k + c] = r.specular[3 * k + c] = 0.0f;
        cart_to_sph(zero, r.normal + 2 * k);
        cart_to_sph(zero, r.wo + 2 * k);
        r.roughness[k] = 1.0f;
        r.delta[k] = 0;
    }
    r.miss[k] = miss ? 1 : 0;
    for (int c = 0; c < 3; ++c) {
        r.L[3 * k + c] = 0.0f;
        r.thp[3 * k + c] = 0.0f;
        if (r.Le) r.Le[3 * k + c] = 0.0f;
    }
}

// The record step of handleIntersections (integrator.cpp:277-296) and
// handleNRCLastBounce (integrator.cpp:193-201)
__global__ void record_scatter_kernel(int64_t n, const dfg_records r, const dfg_shading sd, RecordScatterArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (a.rr_killed && a.rr_killed[i]) return;
    const int32_t pix = a.pixel[i];
    if (pix < 0 || (int64_t)pix >= r.n_pixels) return;
    const uint32_t p = (uint32_t)pix;
    if (!(a.train_enabled && a.depth[i] < a.max_train_depth && training_pixel(p, a.pixel_offset, a.pixel_stride)))
        return;
    if (!(a.alive && a.alive[i])) {
        last_bounce_record(r, sd, i, p, false);
        return;
    }
    // incrementDepth (workqueue.h:26-58)
    const uint32_t d = r.cur_depth[p];
    if (d >= (uint32_t)r.max_depth) return;
    const int64_t k = (int64_t)d * r.n_pixels + p;
    const float c = fabsf(a.cos_theta[i]);
    for (int ch = 0; ch < 3; ++ch) {
        r.L[3 * k + ch] = 0.0f;
        r.thp[3 * k + ch] = a.thp[3 * i + ch];
        r.pos[3 * k + ch] = sd.pos[3 * i + ch];
        r.diffuse[3 * k + ch] = sd.diffuse[3 * i + ch];
        r.specular[3 * k + ch] = sd.specular[3 * i + ch];
        const float b = a.bsdf_val[3 * i + ch];
        r.bsdf_val[3 * k + ch] = a.product_with_cosine ? b * c : b;
        if (r.Le) r.Le[3 * k + ch] = 0.0f;
    }
    cart_to_sph(a.wi_dir + 3 * i, r.dir + 2 * k);
    cart_to_sph(sd.normal + 3 * i, r.normal + 2 * k);
    cart_to_sph(sd.wo + 3 * i, r.wo + 2 * k);
    r.wi_pdf[k] = a.scatter_pdf[i];
    r.mis_weight[k] = a.mis_aware ? a.mis[i] : 1.0f;
    r.delta[k] = a.delta && a.delta[i] ? 1 : 0;
    r.miss[k] = 0;
    r.roughness[k] = sd.roughness[i];
    r.cur_depth[p] = d + 1;
}

// recordRadiance + overwriteLe (workqueue.h:83-106) of handleEmissiveHit and
// handleMiss (integrator.cpp:144-191), with the miss's placeholder record
__global__ void record_radiance_kernel(int64_t n, const dfg_records r, RecordRadianceArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t pix = a.pixel[i];
    if (pix < 0 || (int64_t)pix >= r.n_pixels) return;
    const uint32_t p = (uint32_t)pix;
    if (!(a.train_enabled && training_pixel(p, a.pixel_offset, a.pixel_stride))) return;
    const uint32_t cd = r.cur_depth[p];
    const uint32_t dd = cd < (uint32_t)r.max_depth ? cd : (uint32_t)r.max_depth;
    for (uint32_t d = 0; d < dd; ++d) {
        const int64_t k = (int64_t)d * r.n_pixels + p;
        for (int c = 0; c < 3; ++c) r.L[3 * k + c] = r.L[3 * k + c] + a.contrib[3 * i + c];
    }
    const int last = (int)cd - 1;
    if (r.Le && a.Le && last >= 0 && last < r.max_depth) {
        const int64_t k = (int64_t)last * r.n_pixels + p;
        for (int c = 0; c < 3; ++c) r.Le[3 * k + c] = a.Le[3 * i + c];
    }
    if (a.miss && a.miss[i] && a.depth[i] <= a.max_train_depth) last_bounce_record(r, dfg_shading{}, -1, p, true);
}

}  // namespace

void launch_sample_wi(int64_t n, const SampleWiArgs& a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(sample_wi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a);
}

void launch_record_scatter(int64_t n, const dfg_records& r, const dfg_shading& sd, const RecordScatterArgs& a,
                           hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(record_scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, r, sd, a);
}

void launch_record_radiance(int64_t n, const dfg_records& r, const RecordRadianceArgs& a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(record_radiance_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, r, a);
}

void launch_pdf_image_rows(const float* feat, int D, int rt, int rp, float* rows, float* wi2, hipStream_t s) {
    const int64_t n = (int64_t)rt * rp;
    hipLaunchKernelGGL(pdf_image_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, feat, D, rt, rp,
                       rows, wi2);
}

void launch_directions(int64_t n, const float* wi, float* dir, float* pdf, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(directions_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, wi, dir, pdf);
}

void launch_scatter(int64_t n, float f, const float* wi, const uint8_t* guided, const uint8_t* mix,
                    const float* bsdf_pdf, const float* bsdf_val, const float* cos_theta, float* thp, float* spdf,
                    float* mis, uint8_t* alive, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, f, wi, guided, mix,
                       bsdf_pdf, bsdf_val, cos_theta, thp, spdf, mis, alive);
}

void launch_training_targets(const dfg_records& r, int64_t n_train, uint32_t offset, uint32_t stride,
                             const float* aabb, int mode, float* feat, float* tgt, int2* vidx, int64_t capacity,
                             uint32_t* local, int64_t* block_tot, int64_t* count, hipStream_t s) {
    const unsigned blocks = (unsigned)((n_train + 255) / 256);
    hipLaunchKernelGGL(records_count_kernel, dim3(blocks), dim3(256), 0, s, r, n_train, offset, stride, aabb, mode,
                       local, block_tot);
    hipLaunchKernelGGL(records_scan_kernel, dim3(1), dim3(1024), 0, s, block_tot, (int64_t)blocks, count);
    hipLaunchKernelGGL(records_write_kernel, dim3(blocks), dim3(256), 0, s, r, n_train, offset, stride, aabb, mode,
                       local, block_tot, feat, tgt, vidx, capacity);
}

void launch_nrc_substitute(const dfg_records& r, const int2* vidx, const float* rgb, const int64_t* count,
                           int64_t capacity, hipStream_t s) {
    if (capacity <= 0) return;
    const dim3 grid((unsigned)((capacity + 255) / 256));
    hipLaunchKernelGGL(nrc_apply_kernel, grid, dim3(256), 0, s, r, vidx, rgb, count, capacity);
    hipLaunchKernelGGL(nrc_divide_kernel, grid, dim3(256), 0, s, r, vidx, rgb, count, capacity);
}

}  // namespace dfg
