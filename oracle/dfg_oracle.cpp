// ORACLE — test infrastructure only.  CPU restatement of the reference's
// Distribution-Factorization guiding math (pedrovfigueiredo/guiding-df,
// src/render/guided/{train.h,train.cu,AR.h,learnedDistribution.cpp} and the
// vendored tiny-cuda-nn v1.6 under src/ext/tcnn).  Used only by tests/, the
// smoke check and the bench's CPU-baseline leg, as the checker; never shipped
// and never called by the product path.
//
// Two instantiations of the same templates:
//   *_f64 : everything in double (exp/log from libm) — the FP64 oracle.
//   *_f32 : the reference's fp32 operation order, with the deterministic
//           exp/log restated below, compiled with -ffp-contract=off — the
//           "reference-order" mode used for bit-exact bin/segment checks and
//           as the CPU baseline (OpenMP over rows).  One deliberate departure,
//           shared with the library: grid positions x * scale + 0.5 are
//           evaluated exactly (two fmaf), not rounded to fp32 (pos_fract).
//
// Parity pins (see tests/test_oracle.py): closed forms, invariants, finite
// differences, brute force on tiny grids, numpy matmul chains; the reference
// holds no golden vectors for this path.
//
// Parity unpinned (reference behaviour that only running the reference could
// fix, and the reference cannot run here): tiny-cuda-nn's initial weights for
// a seed (host pcg32 through std::seed_seq, trainer.h:53-58, gpu_matrix.h:291-306,
// and the device generate_random stream, random.h:67-97) — parameters are fed
// to oracle and library alike instead; the rounding of its fp16-accumulating
// WMMA/CUTLASS products (fully_fused_mlp.cu:68,333,436); the CUTLASS split-k
// order; the order of its half2 gradient atomics (grid.h:433-453); the
// outputs of CUDA's expf/logf/__sinf, which the restatement replaces by the
// deterministic exp/log below and libm.
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct {
    int type, dim_begin, n_dims, out_begin, n_out, n_pad, pad_front, param;
    float id_scale, id_offset;
} OrcPart;

typedef struct {
    int in_dims, n_parts;
    OrcPart parts[6];
    int enc_width, n_hidden, n_out, n_out_pad;
    int grid_present, grid_levels, grid_F, grid_interp, grid_type;
    uint32_t grid_offset[25];
    uint32_t grid_res[24];
    float grid_scale[24];
    uint32_t layer_off[10];
    uint32_t n_matrix, n_params, grid_param_begin;
} OrcNet;

}  // extern "C"

namespace {

enum { ENC_IDENTITY = 0, ENC_ONEBLOB = 1, ENC_SH = 2, ENC_GRID = 3, ENC_TRIWAVE = 4, ENC_FREQUENCY = 5 };
enum { INTERP_NEAREST = 0, INTERP_LINEAR = 1, INTERP_SMOOTHSTEP = 2 };
enum { GRID_DENSE = 0, GRID_HASH = 1, GRID_TILED = 2 };

const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Deterministic fp32 exp / log (restatement of the same published steps the
// device code uses: Cody-Waite reduction + Taylor through r^7; fdlibm e_log.c
// reduction f - (f^2/2 - s (f^2/2 + R)) with the atanh series through s^11).
// Correctly rounded IEEE ops only (explicit fmaf / ldexpf where the steps name
// them), no implicit contraction (-ffp-contract=off).
// ---------------------------------------------------------------------------
inline float bits_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
inline float p2(int k) { return bits_f((uint32_t)(k + 127) << 23); }

float det_expf(float x) {
    // rint(x / ln 2) clamped to +-200 (the guards decide outside that range),
    // r = x - k ln2 (hi + lo, two fused multiply-adds), Taylor through r^7 by
    // fused multiply-adds, one correctly rounded ldexp
    const float k = fminf(fmaxf(rintf(x * 1.44269502162933349609f), -200.0f), 200.0f);
    const float r = fmaf(-k, 1.42860676533018704504e-06f, fmaf(-k, 0.693145751953125f, x));
    float p = 1.98412698412698412698e-04f;
    p = fmaf(p, r, 1.38888888888888888889e-03f);
    p = fmaf(p, r, 8.33333333333333333333e-03f);
    p = fmaf(p, r, 4.16666666666666666667e-02f);
    p = fmaf(p, r, 1.66666666666666666667e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.972084045410156f) return 0.0f;
    return ldexpf(p, (int)k);
}

float det_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t b = f_bits(x);
    if (b < 0x00800000u) { x = x * 16777216.0f; b = f_bits(x); e = -24; }
    e += (int)(b >> 23) - 127;
    float m = bits_f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984f) { m = m * 0.5f; e += 1; }
    // fdlibm e_log.c reduction: log(1+f) = f - (hfsq - s (hfsq + R)), s = f/(2+f)
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = 1.81818181818181818182e-01f;  // 2/11
    R = R * z + 2.22222222222222222222e-01f;  // 2/9
    R = R * z + 2.85714285714285714286e-01f;  // 2/7
    R = R * z + 0.4f;                          // 2/5
    R = R * z + 6.66666666666666666667e-01f;  // 2/3
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float ef = (float)e;
    return ef * 0.693145751953125f - ((hfsq - (s * (hfsq + R) + ef * 1.42860676533018704504e-06f)) - f);
}

// IEEE half conversions (round to nearest even) for the fp16 storage steps.
float half_round(float f) {
    uint32_t x = f_bits(f);
    const uint32_t sign = x & 0x80000000u;
    x ^= sign;
    uint16_t h;
    if (x >= 0x7f800000u) {
        h = (uint16_t)(x > 0x7f800000u ? 0x7e00u : 0x7c00u);
    } else if (x >= 0x477ff000u) {  // rounds to >= 65520 -> inf
        h = 0x7c00u;
    } else if (x < 0x38800000u) {   // subnormal half
        const float a = bits_f(x);
        const float scaled = a * 16777216.0f;  // 2^24: half subnormal ulp is 2^-24
        h = (uint16_t)rintf(scaled);
    } else {
        const uint32_t mant = x & 0x7fffffu;
        const uint32_t exp = (x >> 23) - 112;
        uint32_t v = (exp << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (v & 1u))) v += 1;
        h = (uint16_t)v;
    }
    // back to float
    const uint32_t hs = (uint32_t)h;
    const uint32_t he = (hs >> 10) & 0x1fu, hm = hs & 0x3ffu;
    float r;
    if (he == 0) r = (float)hm * (1.0f / 16777216.0f);
    else if (he == 31) r = hm ? NAN : INFINITY;
    else r = bits_f(((he + 112) << 23) | (hm << 13));
    return sign ? -r : r;
}

template <typename R> inline R r_exp(R x);
template <> inline double r_exp<double>(double x) { return std::exp(x); }
template <> inline float r_exp<float>(float x) { return det_expf(x); }
template <typename R> inline R r_log(R x);
template <> inline double r_log<double>(double x) { return std::log(x); }
template <> inline float r_log<float>(float x) { return det_logf(x); }

// ---------------------------------------------------------------------------
// Encodings (tcnn encodings/*.h).  R = double | float.
// ---------------------------------------------------------------------------
template <typename R>
R quartic_cdf(R x, R inv_radius) {  // common_device.h:492-497
    const R u = x * inv_radius;
    const R u2 = u * u;
    const R u4 = u2 * u2;
    const R c23 = (R)2 / (R)3, c15 = (R)1 / (R)5;
    R v = (R)15 / (R)16 * u * ((R)1 - c23 * u2 + c15 * u4) + (R)0.5;
    return v < 0 ? (R)0 : (v > 1 ? (R)1 : v);
}

// grid_index, grid.h:164-186, CoherentPrime hash; uint32 stride as in the reference
uint32_t grid_index3(int type, uint32_t hsize, uint32_t res, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t pg[3] = {x, y, z};
    uint32_t stride = 1, index = 0;
    for (int d = 0; d < 3 && stride <= hsize; ++d) {
        index += pg[d] * stride;
        stride *= res;
    }
    if (type == GRID_HASH && hsize < stride) index = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
    return index % hsize;
}

// floor(v) as an int for every value: NaN -> 0, beyond the int range
// saturated (the device's conversion; the plain cast would be undefined).
template <typename R>
int floor_to_int(R v) {
    const R f = std::floor(v);
    if (!(f == f)) return 0;
    return f >= (R)2147483648.0 ? 2147483647 : (f < (R)-2147483648.0 ? (-2147483647 - 1) : (int)f);
}

// common_device.h:418-431.  The position p = in * scale + 0.5 is taken
// exactly in both modes (the reference rounds it to fp32, which costs up to
// half an ulp, 2^-6 at scale 2^19, of the fine levels' weights):
//   double: the product of two floats and the + 0.5 are exact in double
//           wherever p is near an integer, so floor and p - floor are exact;
//   float:  t = floor(fmaf(in, scale, 0.5)), f = fmaf(in, scale, 0.5 - t), and
//           t - 1 when f < 0 (the first fmaf rounded up onto an integer): the
//           cell is floor(p) and f is p - cell rounded once.
template <typename R>
void pos_fract(R in, R scale, int interp, R& frac, uint32_t& cell);

template <>
void pos_fract<double>(double in, double scale, int interp, double& frac, uint32_t& cell) {
    double p = in * scale + 0.5;
    const double t = std::floor(p);
    cell = (uint32_t)floor_to_int(t);
    p = p - t;
    if (interp == INTERP_SMOOTHSTEP) p = p * p * (3.0 - 2.0 * p);
    frac = p;
}

template <>
void pos_fract<float>(float in, float scale, int interp, float& frac, uint32_t& cell) {
    float t = std::floor(fmaf(in, scale, 0.5f));
    float p = fmaf(in, scale, 0.5f - t);
    if (p < 0.0f) {
        t = t - 1.0f;
        p = fmaf(in, scale, 0.5f - t);
    }
    cell = (uint32_t)floor_to_int(t);
    if (interp == INTERP_SMOOTHSTEP) p = p * p * (3.0f - 2.0f * p);
    frac = p;
}

template <typename R>
struct Corners {
    uint32_t idx[8];
    R w[8];
};

template <typename R>
void grid_corners(const OrcNet& nt, int level, const R* pos, Corners<R>& c) {
    const R scale = (R)nt.grid_scale[level];
    const uint32_t res = nt.grid_res[level];
    const uint32_t hsize = nt.grid_offset[level + 1] - nt.grid_offset[level];
    R f[3];
    uint32_t cell[3];
    for (int d = 0; d < 3; ++d) pos_fract<R>(pos[d], scale, nt.grid_interp == INTERP_NEAREST ? INTERP_LINEAR : nt.grid_interp, f[d], cell[d]);
    if (nt.grid_interp == INTERP_NEAREST) {
        for (int k = 0; k < 8; ++k) { c.idx[k] = 0; c.w[k] = 0; }
        c.idx[0] = grid_index3(nt.grid_type, hsize, res, cell[0], cell[1], cell[2]);
        c.w[0] = 1;
        return;
    }
    for (int k = 0; k < 8; ++k) {
        R w = 1;
        uint32_t p[3];
        for (int d = 0; d < 3; ++d) {
            if (k & (1 << d)) { w *= f[d]; p[d] = cell[d] + 1; }
            else { w *= (R)1 - f[d]; p[d] = cell[d]; }
        }
        c.idx[k] = grid_index3(nt.grid_type, hsize, res, p[0], p[1], p[2]);
        c.w[k] = w;
    }
}

template <typename R>
void sh4(int degree, R ix, R iy, R iz, R* o) {  // spherical_harmonics.h:46-152 (closed forms)
    const R x = ix * 2 - 1, y = iy * 2 - 1, z = iz * 2 - 1;
    const R xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    const R s = std::sqrt((R)kPi);
    o[0] = (R)1 / (2 * s);
    if (degree <= 1) return;
    o[1] = -std::sqrt((R)3) * y / (2 * s);
    o[2] = std::sqrt((R)3) * z / (2 * s);
    o[3] = -std::sqrt((R)3) * x / (2 * s);
    if (degree <= 2) return;
    o[4] = std::sqrt((R)15) * xy / (2 * s);
    o[5] = -std::sqrt((R)15) * yz / (2 * s);
    o[6] = std::sqrt((R)5) * (3 * z2 - 1) / (4 * s);
    o[7] = -std::sqrt((R)15) * xz / (2 * s);
    o[8] = std::sqrt((R)15) * (x2 - y2) / (4 * s);
    if (degree <= 3) return;
    o[9] = std::sqrt((R)70) * y * (-3 * x2 + y2) / (8 * s);
    o[10] = std::sqrt((R)105) * xy * z / (2 * s);
    o[11] = std::sqrt((R)42) * y * (1 - 5 * z2) / (8 * s);
    o[12] = std::sqrt((R)7) * z * (5 * z2 - 3) / (4 * s);
    o[13] = std::sqrt((R)42) * x * (1 - 5 * z2) / (8 * s);
    o[14] = std::sqrt((R)105) * z * (x2 - y2) / (4 * s);
    o[15] = std::sqrt((R)70) * x * (-x2 + 3 * y2) / (8 * s);
}

// Encode one row: in[in_dims] -> out[enc_width].  `table` is the grid part of
// the parameter vector (R values), or null.
template <typename R>
void encode_row(const OrcNet& nt, const R* in, const R* table, R* out) {
    for (int pi = 0; pi < nt.n_parts; ++pi) {
        const OrcPart& P = nt.parts[pi];
        const int ob = P.out_begin + (P.pad_front ? P.n_pad : 0);
        const int pb = P.pad_front ? P.out_begin : P.out_begin + P.n_out;
        for (int j = 0; j < P.n_pad; ++j) out[pb + j] = 1;
        const R* x = in + P.dim_begin;
        switch (P.type) {
            case ENC_IDENTITY:  // identity.h:46-69
                for (int d = 0; d < P.n_dims; ++d) out[ob + d] = x[d] * (R)P.id_scale + (R)P.id_offset;
                break;
            case ENC_ONEBLOB: {  // oneblob.h:98-127, wraparound over {-1, 0, +1}
                const int nb = P.param;
                for (int d = 0; d < P.n_dims; ++d) {
                    for (int k = 0; k < nb; ++k) {
                        const R lb = (R)k / nb, rb = (R)(k + 1) / nb;
                        R left = 0, right = 0;
                        for (int sh = -1; sh <= 1; ++sh) {
                            left += quartic_cdf<R>(lb - x[d] + sh, (R)nb);
                            right += quartic_cdf<R>(rb - x[d] + sh, (R)nb);
                        }
                        out[ob + d * nb + k] = right - left;
                    }
                }
                break;
            }
            case ENC_SH:
                sh4<R>(P.param, x[0], x[1], x[2], out + ob);
                break;
            case ENC_TRIWAVE: {  // triangle_wave.h:46-82
                const int nf = P.param;
                for (int d = 0; d < P.n_dims; ++d)
                    for (int f = 0; f < nf; ++f) {
                        const R v = std::ldexp(x[d], f - 1) + (R)f * (R)0.25;
                        out[ob + d * nf + f] = std::fabs(v - std::floor(v) - (R)0.5) * 4 - 1;
                    }
                break;
            }
            case ENC_FREQUENCY: {  // frequency.h:46-94: sin(2^f pi x + (j%2) pi/2)
                const int nf = P.param;
                for (int d = 0; d < P.n_dims; ++d)
                    for (int j = 0; j < 2 * nf; ++j) {
                        const double arg = std::ldexp((double)x[d], j / 2) * kPi + (j % 2) * kPi / 2;
                        out[ob + d * 2 * nf + j] = (R)std::sin(arg);
                    }
                break;
            }
            case ENC_GRID: {  // grid.h:219-392
                const int F = nt.grid_F;
                for (int l = 0; l < nt.grid_levels; ++l) {
                    Corners<R> c;
                    grid_corners<R>(nt, l, x, c);
                    const R* lt = table + (size_t)nt.grid_offset[l] * F;
                    for (int f = 0; f < F; ++f) {
                        R acc = 0;
                        const int nk = nt.grid_interp == INTERP_NEAREST ? 1 : 8;
                        for (int k = 0; k < nk; ++k) acc += c.w[k] * lt[(size_t)c.idx[k] * F + f];
                        out[ob + l * F + f] = acc;
                    }
                }
                break;
            }
        }
    }
}

// MLP forward for one row; keeps the post-ReLU activations when acts != null
// (acts[(l) * 64 + j] = a_{l+1}[j]).
template <typename R>
void mlp_row(const OrcNet& nt, const R* params, const R* x, R* out, R* acts) {
    R a[64], b[64];
    const int H = nt.n_hidden;
    const R* W0 = params + nt.layer_off[0];
    for (int m = 0; m < 64; ++m) {
        R s = 0;
        for (int k = 0; k < nt.enc_width; ++k) s += W0[m * nt.enc_width + k] * x[k];
        a[m] = s > 0 ? s : 0;
    }
    if (acts) for (int m = 0; m < 64; ++m) acts[m] = a[m];
    for (int l = 1; l < H; ++l) {
        const R* W = params + nt.layer_off[l];
        for (int m = 0; m < 64; ++m) {
            R s = 0;
            for (int k = 0; k < 64; ++k) s += W[m * 64 + k] * a[k];
            b[m] = s > 0 ? s : 0;
        }
        memcpy(a, b, sizeof(a));
        if (acts) for (int m = 0; m < 64; ++m) acts[l * 64 + m] = a[m];
    }
    const R* WH = params + nt.layer_off[H];
    for (int m = 0; m < nt.n_out_pad; ++m) {
        R s = 0;
        for (int k = 0; k < 64; ++k) s += WH[m * 64 + k] * a[k];
        out[m] = s;
    }
}

// ---------------------------------------------------------------------------
// DF head (train.h:345-520, 649-681, 793-812, 844-893)
// ---------------------------------------------------------------------------
// max initialised to 0 (train.h:846).  N exp(z_i - m) / sum: the reference
// divides through a second exponential, exp(z_i - m - log sum); the product
// with N / sum is the same quantity (exact in FP64, within a few ulp in FP32)
// and is what the library computes, so the FP32 mode stays its bit-exact twin.
template <typename R>
void softmax_times_n(const R* z, int N, R* p) {
    R mx = 0;
    for (int i = 0; i < N; ++i) mx = std::fmax(mx, z[i]);
    R s = 0;
    for (int i = 0; i < N; ++i) {
        p[i] = r_exp<R>(z[i] - mx);
        s += p[i];
    }
    const R scale = (R)N / s;
    for (int i = 0; i < N; ++i) p[i] = s > 0 ? p[i] * scale : (R)INFINITY;
}

template <typename R>
struct Seg { int i; R lo, hi; };

template <typename R>
Seg<R> search_linear(const R* p, int N, R u) {  // train.h:361-397
    Seg<R> s;
    R c = p[0] * (R)0.5 / (R)N;
    s.lo = 0; s.hi = c; s.i = 0;
    if (u < c) return s;
    R prev = c;
    for (int i = 1; i < N; ++i) {
        c = prev + (p[i - 1] + p[i]) / (R)(2 * N);
        if (u < c) { s.i = i; s.lo = prev; s.hi = c; return s; }
        prev = c;
    }
    s.i = N; s.lo = prev; s.hi = prev + p[N - 1] * (R)0.5 / (R)N;
    return s;
}

template <typename R>
Seg<R> search_nearest(const R* p, int N, R u) {  // train.h:345-359
    Seg<R> s;
    R prev = 0, c = 0;
    for (int i = 0; i < N; ++i) {
        c = prev + p[i] / (R)N;
        if (u < c) { s.i = i; s.lo = prev; s.hi = c; return s; }
        prev = c;
    }
    s.i = N; s.lo = prev; s.hi = c;
    return s;
}

template <typename R>
R sample_linear(const R* p, int N, const Seg<R>& sg, R u, R u_extra) {  // train.h:453-519
    const R eps = (R)1e-5f;
    const R invN = (R)1 / (R)N;
    int vi = sg.i, vim1 = vi - 1;
    R v = (R)vim1 + (R)0.5;
    R ecm = 1, voff = 0;
    if (vim1 < 0) { vim1 = vi; voff = (R)0.5; ecm = (R)0.5; }
    else if (vi > N - 1) { vi = vim1; ecm = (R)0.5; }
    const R m = (p[vi] - p[vim1]) * (R)N;
    const R B = p[vim1] * ((R)1 + v) - p[vi] * v;
    v += voff;
    const R A = m * (R)0.5;
    R x;
    if (std::fabs(A) < eps) {
        const R vd = u - sg.lo, cd = sg.hi - sg.lo;
        const R w = cd > eps ? vd / cd : u_extra;
        x = (v + w * ecm) * invN;
    } else {
        // branch on the reference's global discriminant (train.h:494-497);
        // root in the segment's coordinate t = x - v/N, where the cdf is
        // lo + p0 t + A t^2 (p0 = p[vim1]) and the selected root is
        // 2 d / (p0 + sqrt(p0^2 + 4 A d)), d = u - lo, for either sign of A
        // (the same root; no cancellation of terms of size A (v/N)^2)
        const R C = -(m * (R)0.5) * (v * v * invN * invN) - (B * v * invN) + sg.lo - u;
        const R delta = B * B - (R)4 * A * C;
        if (delta < eps) {
            x = (v + u_extra * ecm) * invN;
        } else {
            const R p0 = p[vim1];
            const R d = u - sg.lo;
            const R disc = std::fmax(p0 * p0 + (R)4 * A * d, (R)0);
            const R den = p0 + std::sqrt(disc);
            const R t = den > 0 ? ((R)2 * d) / den : (R)0;
            x = v * invN + t;
        }
    }
    return std::fmin(x, (R)1 - eps);
}

template <typename R>
R sample_nearest(int N, const Seg<R>& sg, R u, R u_extra) {
    const R eps = (R)1e-5f;
    const R vd = u - sg.lo, cd = sg.hi - sg.lo;
    const R w = cd > eps ? vd / cd : u_extra;
    return std::fmin(((R)sg.i + w) / (R)N, (R)1 - eps);
}

// sample_pdf, train.h:654-681.  The reference's indices on [0, 1] and above;
// below the first bin centre's half-bin its i1 runs negative (an out-of-bounds
// read), here i1 stays 0; NaN x gives a NaN pdf.
template <typename R>
R pdf_linear(const R* p, int N, R x) {
    R idx = x * (R)N;
    idx -= (R)0.5;
    int i0 = std::min(floor_to_int(idx), N - 1);
    int i1 = i0 + 1;
    R w = idx - (R)i0;
    if (i0 < 0) { i0 = N - 1; w = 1; }
    if (i1 > N - 1) { i1 = 0; w = 0; }
    if (i1 < 0) i1 = 0;
    return ((R)1 - w) * p[i0] + w * p[i1];
}

template <typename R>
R pdf_nearest(const R* p, int N, R x) {
    int i = std::min(floor_to_int(x * (R)N), N - 1);
    if (i < 0) i = 0;
    return p[i];
}

template <typename R>
R theta_input(R th, int N, int linear) {  // add_theta_to_input_F, train.h:793-812
    const R t = th * (R)N;
    const int fi = std::min(floor_to_int(t), N - 1);
    if (linear) return ((R)fi + (t - (R)fi)) / (R)N;
    return (R)fi / (R)N;
}

template <typename R>
R head(const R* z, int N, int linear, bool sample, R u, R ue, R x_given, R* pdf, R* bins, int* seg) {
    std::vector<R> p(N);
    softmax_times_n<R>(z, N, p.data());
    if (bins) for (int i = 0; i < N; ++i) bins[i] = p[i];
    R x = x_given;
    if (seg) *seg = -1;
    if (sample) {
        if (linear) { Seg<R> s = search_linear<R>(p.data(), N, u); if (seg) *seg = s.i; x = sample_linear<R>(p.data(), N, s, u, ue); }
        else { Seg<R> s = search_nearest<R>(p.data(), N, u); if (seg) *seg = s.i; x = sample_nearest<R>(N, s, u, ue); }
    }
    *pdf = linear ? pdf_linear<R>(p.data(), N, x) : pdf_nearest<R>(p.data(), N, x);
    return x;
}

// Whole query of one row: theta net -> head -> phi net -> head (-> NRC).
template <typename R>
void query_row(const OrcNet& th, const OrcNet& ph, const R* pth, const R* pph, const float* feat, int D,
               const float* u, int guided, const float* wi_in, int linear, float* wi_out, R* extra_logits) {
    R in[17], x[128], lt[64], lp[64];
    for (int d = 0; d < D; ++d) in[d] = (R)feat[d];
    encode_row<R>(th, in, pth + th.grid_param_begin, x);
    mlp_row<R>(th, pth, x, lt, nullptr);
    R pt, pp;
    const R u0 = head<R>(lt, th.n_out, linear, guided != 0, (R)u[0], (R)u[2], (R)wi_in[0], &pt, nullptr, nullptr);
    in[D] = theta_input<R>(u0, th.n_out, linear);
    encode_row<R>(ph, in, pph + ph.grid_param_begin, x);
    mlp_row<R>(ph, pph, x, lp, nullptr);
    const R u1 = head<R>(lp, ph.n_out, linear, guided != 0, (R)u[1], (R)u[3], (R)wi_in[1], &pp, nullptr, nullptr);
    wi_out[0] = (float)u0;
    wi_out[1] = (float)u1;
    wi_out[2] = (float)((R)wi_in[2] * pt);
    wi_out[3] = (float)((R)wi_in[3] * pp);
    if (extra_logits) {
        for (int i = 0; i < th.n_out; ++i) extra_logits[i] = lt[i];
        for (int i = 0; i < ph.n_out; ++i) extra_logits[64 + i] = lp[i];
    }
}

template <typename R>
std::vector<R> to_r(const float* p, size_t n) {
    std::vector<R> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (R)p[i];
    return v;
}

}  // namespace

extern "C" {

// --- primitives ----------------------------------------------------------------
float orc_det_expf(float x) { return det_expf(x); }
float orc_det_logf(float x) { return det_logf(x); }
float orc_half_round(float x) { return half_round(x); }

void orc_grid_index(const OrcNet* nt, const float* in, int64_t n, int stride, int pos_dim, uint32_t* idx) {
    for (int64_t i = 0; i < n; ++i) {
        for (int l = 0; l < nt->grid_levels; ++l) {
            Corners<float> c;
            grid_corners<float>(*nt, l, in + i * stride + pos_dim, c);
            for (int k = 0; k < 8; ++k) idx[(i * nt->grid_levels + l) * 8 + k] = c.idx[k];
        }
    }
}

uint32_t orc_grid_index3(int type, uint32_t hsize, uint32_t res, uint32_t x, uint32_t y, uint32_t z) {
    return grid_index3(type, hsize, res, x, y, z);
}

void orc_encode_f64(const OrcNet* nt, const double* params, const double* in, int64_t n, int stride, double* out) {
    for (int64_t i = 0; i < n; ++i) encode_row<double>(*nt, in + i * stride, params + nt->grid_param_begin, out + i * nt->enc_width);
}

void orc_mlp_f64(const OrcNet* nt, const double* params, const double* x, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) mlp_row<double>(*nt, params, x + i * nt->enc_width, out + i * nt->n_out_pad, nullptr);
}

// Stateless head on given logits.  mode bit 0: linear; f64 or f32 variant.
void orc_head_f64(const double* logits, int64_t n, int N, int linear, const double* u, const double* ue,
                  const uint8_t* sample, const double* x_given, double* x_out, double* pdf, double* bins, int* seg) {
    for (int64_t i = 0; i < n; ++i)
        x_out[i] = head<double>(logits + i * N, N, linear, sample ? sample[i] != 0 : true, u[i], ue[i],
                                x_given ? x_given[i] : 0.0, pdf + i, bins ? bins + i * N : nullptr, seg ? seg + i : nullptr);
}

void orc_head_f32(const float* logits, int64_t n, int N, int linear, const float* u, const float* ue,
                  const uint8_t* sample, const float* x_given, float* x_out, float* pdf, float* bins, int* seg) {
    for (int64_t i = 0; i < n; ++i)
        x_out[i] = head<float>(logits + i * N, N, linear, sample ? sample[i] != 0 : true, u[i], ue[i],
                               x_given ? x_given[i] : 0.0f, pdf + i, bins ? bins + i * N : nullptr, seg ? seg + i : nullptr);
}

double orc_pdf_linear_f64(const double* p, int N, double x) { return pdf_linear<double>(p, N, x); }
double orc_theta_input_f64(double th, int N, int linear) { return theta_input<double>(th, N, linear); }
float orc_theta_input_f32(float th, int N, int linear) { return theta_input<float>(th, N, linear); }

// Full query.  params are the fp32 values of the weights the GPU uses.
void orc_query_f64(const OrcNet* th, const OrcNet* ph, const float* pth, const float* pph, const float* feat,
                   int stride, int D, int64_t n, const float* u, const uint8_t* guided, const float* wi_in, int linear,
                   float* wi_out, double* logits /*[n][128] or null*/) {
    const std::vector<double> a = to_r<double>(pth, th->n_params), b = to_r<double>(pph, ph->n_params);
    for (int64_t i = 0; i < n; ++i)
        query_row<double>(*th, *ph, a.data(), b.data(), feat + i * stride, D, u + 4 * i, guided ? guided[i] : 1,
                          wi_in + 4 * i, linear, wi_out + 4 * i, logits ? logits + i * 128 : nullptr);
}

// FP32 reference-order full query with OpenMP over rows: the CPU baseline.
int orc_query_f32_omp(const OrcNet* th, const OrcNet* ph, const float* pth, const float* pph, const float* feat,
                      int stride, int D, int64_t n, const float* u, const uint8_t* guided, const float* wi_in,
                      int linear, float* wi_out) {
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        query_row<float>(*th, *ph, pth, pph, feat + i * stride, D, u + 4 * i, guided ? guided[i] : 1, wi_in + 4 * i,
                         linear, wi_out + 4 * i, nullptr);
    return threads;
}

// FP32 reference-order NRC forward with OpenMP over rows (CPU baseline leg).
void orc_nrc_f32_omp(const OrcNet* nt, const float* params, const float* feat, int stride, int64_t n, float* rgb) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float in[16], x[128], o[64];
        for (int d = 0; d < 15; ++d) in[d] = feat[i * stride + d];
        encode_row<float>(*nt, in, params + nt->grid_param_begin, x);
        mlp_row<float>(*nt, params, x, o, nullptr);
        for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = o[c];
    }
}

// NRC forward (f64): feat[n][15] -> rgb[n][3]
void orc_nrc_f64(const OrcNet* nt, const float* params, const float* feat, int stride, int64_t n, double* rgb) {
    const std::vector<double> p = to_r<double>(params, nt->n_params);
    double in[16], x[128], o[64];
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 15; ++d) in[d] = feat[i * stride + d];
        encode_row<double>(*nt, in, p.data() + nt->grid_param_begin, x);
        mlp_row<double>(*nt, p.data(), x, o, nullptr);
        for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = o[c];
    }
}

// ---------------------------------------------------------------------------
// Training oracle (FP64): loss and full analytic parameter gradient.
// DF: combine_AR_pdf_and_dl_doutput (train.h:931-965), AR.h:50-116.
// Returns the mean loss; grads (if non-null) receive d(mean loss)/d(params).
// ---------------------------------------------------------------------------
static void mlp_backward(const OrcNet& nt, const double* P, const double* x, const double* acts, const double* gout,
                         double* dP, double* dx) {
    const int H = nt.n_hidden;
    double g[64], ga[64];
    // output layer
    const double* WH = P + nt.layer_off[H];
    double* dWH = dP + nt.layer_off[H];
    const double* aH = acts + (H - 1) * 64;
    for (int m = 0; m < nt.n_out_pad; ++m)
        for (int k = 0; k < 64; ++k) dWH[m * 64 + k] += gout[m] * aH[k];
    for (int k = 0; k < 64; ++k) {
        double s = 0;
        for (int m = 0; m < nt.n_out_pad; ++m) s += WH[m * 64 + k] * gout[m];
        g[k] = aH[k] > 0 ? s : 0;
    }
    for (int l = H - 1; l >= 1; --l) {
        const double* W = P + nt.layer_off[l];
        double* dW = dP + nt.layer_off[l];
        const double* al = acts + (l - 1) * 64;
        for (int m = 0; m < 64; ++m)
            for (int k = 0; k < 64; ++k) dW[m * 64 + k] += g[m] * al[k];
        for (int k = 0; k < 64; ++k) {
            double s = 0;
            for (int m = 0; m < 64; ++m) s += W[m * 64 + k] * g[m];
            ga[k] = al[k] > 0 ? s : 0;
        }
        memcpy(g, ga, sizeof(g));
    }
    const double* W0 = P + nt.layer_off[0];
    double* dW0 = dP + nt.layer_off[0];
    for (int m = 0; m < 64; ++m)
        for (int k = 0; k < nt.enc_width; ++k) dW0[m * nt.enc_width + k] += g[m] * x[k];
    for (int k = 0; k < nt.enc_width; ++k) {
        double s = 0;
        for (int m = 0; m < 64; ++m) s += W0[m * nt.enc_width + k] * g[m];
        dx[k] = s;
    }
}

static void grid_backward(const OrcNet& nt, const double* pos, const double* dx, double* dP) {
    int ob = -1;
    for (int p = 0; p < nt.n_parts; ++p)
        if (nt.parts[p].type == ENC_GRID) ob = nt.parts[p].out_begin;
    if (ob < 0) return;
    const int F = nt.grid_F;
    for (int l = 0; l < nt.grid_levels; ++l) {
        Corners<double> c;
        grid_corners<double>(nt, l, pos, c);
        double* gt = dP + nt.grid_param_begin + (size_t)nt.grid_offset[l] * F;
        const int nk = nt.grid_interp == INTERP_NEAREST ? 1 : 8;
        for (int k = 0; k < nk; ++k)
            for (int f = 0; f < F; ++f) gt[(size_t)c.idx[k] * F + f] += c.w[k] * dx[ob + l * F + f];
    }
}

static int grid_dim(const OrcNet& nt) {
    for (int p = 0; p < nt.n_parts; ++p)
        if (nt.parts[p].type == ENC_GRID) return nt.parts[p].dim_begin;
    return 0;
}

// d pdf(x) / d z_j for bins p = N softmax(z) (linear or nearest interpolation)
static void dpdf_dz(const double* p, int N, double x, int linear, double* out) {
    const double inv = 1.0 / N;
    if (linear) {
        const double idx = x * N - 0.5;
        int i0 = std::min(floor_to_int(idx), N - 1), i1 = i0 + 1;
        double w = idx - i0;
        if (i0 < 0) { i0 = N - 1; w = 1; }
        if (i1 > N - 1) { i1 = 0; w = 0; }
        if (i1 < 0) i1 = 0;
        const double s0 = p[i0] * inv, s1 = p[i1] * inv;
        for (int j = 0; j < N; ++j) {
            const double sj = p[j] * inv;
            out[j] = N * ((1 - w) * s0 * ((j == i0) - sj) + w * s1 * ((j == i1) - sj));
        }
    } else {
        int i0 = std::min(floor_to_int(x * N), N - 1);
        if (i0 < 0) i0 = 0;
        const double s0 = p[i0] * inv;
        for (int j = 0; j < N; ++j) out[j] = N * s0 * ((j == i0) - p[j] * inv);
    }
}

// half_bins: store the softmax in fp16 before the xN like the reference's
// training forward (softmaxMultFactor with precision_t = __half, train.h:874-893).
double orc_df_train_f64(const OrcNet* th, const OrcNet* ph, const double* pth, const double* pph, const float* feat,
                        int stride, int D, const float* target, int64_t B, int linear, int half_bins, double* gth,
                        double* gph) {
    const double eps = (double)1e-5f;
    const double inv4pi = (double)0.07957747154594766788f;
    if (gth) memset(gth, 0, sizeof(double) * th->n_params);
    if (gph) memset(gph, 0, sizeof(double) * ph->n_params);
    double total = 0;
    std::vector<double> in(17), xt(128), xp(128), at(64 * 8), ap(64 * 8), lt(64), lp(64), bt(64), bp(64), dt(64), dp(64),
        got(64), gop(64), dxt(128), dxp(128);
    for (int64_t b = 0; b < B; ++b) {
        const float* fr = feat + b * stride;
        const float* t = target + b * 9;
        for (int d = 0; d < D; ++d) in[d] = fr[d];
        encode_row<double>(*th, in.data(), pth + th->grid_param_begin, xt.data());
        mlp_row<double>(*th, pth, xt.data(), lt.data(), at.data());
        in[D] = theta_input<double>((double)t[0], th->n_out, linear);
        encode_row<double>(*ph, in.data(), pph + ph->grid_param_begin, xp.data());
        mlp_row<double>(*ph, pph, xp.data(), lp.data(), ap.data());
        softmax_times_n<double>(lt.data(), th->n_out, bt.data());
        softmax_times_n<double>(lp.data(), ph->n_out, bp.data());
        if (half_bins) {
            for (int j = 0; j < th->n_out; ++j) bt[j] = (double)half_round((float)(bt[j] / th->n_out)) * th->n_out;
            for (int j = 0; j < ph->n_out; ++j) bp[j] = (double)half_round((float)(bp[j] / ph->n_out)) * ph->n_out;
        }
        const double x0 = t[0], x1 = t[1];
        const double pt = linear ? pdf_linear<double>(bt.data(), th->n_out, x0) : pdf_nearest<double>(bt.data(), th->n_out, x0);
        const double pp = linear ? pdf_linear<double>(bp.data(), ph->n_out, x1) : pdf_nearest<double>(bp.data(), ph->n_out, x1);
        const double Li = ((double)t[2] + t[3] + t[4]) / 3.0;
        const double wiPdf = (double)t[5] + eps;
        const double q = ((std::fabs(pt) * std::fabs(pp)) * inv4pi) * (1.0 / inv4pi) + eps;
        total += -Li / wiPdf * std::log(q);
        if (!gth && !gph) continue;
        // d(-Li/wiPdf log q)/dz / B
        const double pre = -Li / wiPdf / q / (double)B;
        dpdf_dz(bt.data(), th->n_out, x0, linear, dt.data());
        dpdf_dz(bp.data(), ph->n_out, x1, linear, dp.data());
        // a loss-head gradient that is not finite contributes nothing (the
        // library's rule; the reference's NaN would reach every weight)
        auto fin = [](double g) { return std::fabs(g) <= 1.7976931348623157e308 ? g : 0.0; };
        for (int j = 0; j < th->n_out_pad; ++j) got[j] = j < th->n_out ? fin(pre * std::fabs(pp) * dt[j]) : 0;
        for (int j = 0; j < ph->n_out_pad; ++j) gop[j] = j < ph->n_out ? fin(pre * std::fabs(pt) * dp[j]) : 0;
        if (gth) {
            mlp_backward(*th, pth, xt.data(), at.data(), got.data(), gth, dxt.data());
            grid_backward(*th, in.data() + grid_dim(*th), dxt.data(), gth);
        }
        if (gph) {
            mlp_backward(*ph, pph, xp.data(), ap.data(), gop.data(), gph, dxp.data());
            grid_backward(*ph, in.data() + grid_dim(*ph), dxp.data(), gph);
        }
    }
    return total / (double)B;
}

// NRC RelativeL2Luminance (relative_l2_luminance.h:39-88): returns the summed
// loss / B (what nrc.cpp:166 reports); grads = d(sum of values)/d(params)
// with the luminance treated as a constant.
double orc_nrc_train_f64(const OrcNet* nt, const double* P, const float* feat, int stride, const float* target,
                         int64_t B, double* grad) {
    if (grad) memset(grad, 0, sizeof(double) * nt->n_params);
    double total = 0;
    std::vector<double> in(16), x(128), a(64 * 8), o(64), go(64), dx(128);
    const double n_total = (double)B * 3.0;
    for (int64_t b = 0; b < B; ++b) {
        for (int d = 0; d < 15; ++d) in[d] = feat[b * stride + d];
        encode_row<double>(*nt, in.data(), P + nt->grid_param_begin, x.data());
        mlp_row<double>(*nt, P, x.data(), o.data(), a.data());
        const double lum = 0.299 * o[0] + 0.587 * o[1] + 0.114 * o[2];
        const double den = lum * lum + 0.01;
        for (int c = 0; c < nt->n_out_pad; ++c) go[c] = 0;
        for (int c = 0; c < 3; ++c) {
            const double diff = o[c] - target[b * 3 + c];
            total += diff * diff / den / n_total;
            const double gc = 2 * diff / den / n_total;
            go[c] = std::fabs(gc) <= 1.7976931348623157e308 ? gc : 0.0;  // the loss-head rule of the library
        }
        if (grad) mlp_backward(*nt, P, x.data(), a.data(), go.data(), grad, dx.data());
    }
    return total / (double)B;
}

// Adam (adam.h:47-120) + EMA (ema.h:62-77) on one parameter vector, FP64.
void orc_adam_ema_f64(int64_t n, int64_t n_matrix, double lr, double b1, double b2, double eps, double l2,
                      double ema_decay, int global_step, const double* grad /*unscaled*/, double* master,
                      double* m, double* v, uint32_t* steps, const double* w_lowp /*fp16 weights after cast*/,
                      double* ema) {
    const double dbo = 1.0 - std::pow(ema_decay, global_step - 1);
    const double dbn = 1.0 / (1.0 - std::pow(ema_decay, global_step));
    for (int64_t i = 0; i < n; ++i) {
        double g = grad[i];
        const bool matrix = i < n_matrix;
        if (matrix || g != 0) {
            if (matrix) g += l2 * master[i];
            m[i] = b1 * m[i] + (1 - b1) * g;
            v[i] = b2 * v[i] + (1 - b2) * g * g;
            const uint32_t st = ++steps[i];
            const double lr_t = lr * std::sqrt(1 - std::pow(b2, st)) / (1 - std::pow(b1, st));
            const double elr = std::max(lr_t / (std::sqrt(v[i]) + eps), 0.0);
            master[i] = master[i] - elr * m[i];
        }
        if (ema && w_lowp) ema[i] = (ema[i] * ema_decay * dbo + w_lowp[i] * (1 - ema_decay)) * dbn;
    }
}

}  // extern "C"
