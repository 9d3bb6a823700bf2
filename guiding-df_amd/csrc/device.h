// Device-side building blocks of the guide on CDNA4 (gfx950):
//   * deterministic exp/log used by the softmax (so bins are reproducible),
//   * the input encodings (grid, spherical harmonics, OneBlob, triangle wave,
//     frequency, identity) evaluated one row per lane,
//   * the Distribution-Factorization head: softmax x N, sequential CDF search,
//     piecewise-linear inverse CDF, pdf evaluation, loss gradient,
//   * 64-wide MLP layers on v_mfma_f32_16x16x32_f16 with fp32 accumulation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "spec.h"

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace dfg {

#define DFG_EPS 1e-5f                 // M_EPSILON, krrmath/constants.h:19
#define DFG_INV_4PI 0.07957747154594766788f  // M_INV_4PI, krrmath/constants.h:17

// ---------------------------------------------------------------------------
// Deterministic exp / log.  Both use only correctly rounded IEEE-754
// operations (+, -, *, /, explicit fused multiply-add, ldexp) and exact bit
// manipulation, with implicit FMA contraction disabled, so the host
// restatement of the same steps produces bit-identical results.
//   exp: Cody & Waite additive range reduction x = k ln2 + r with ln2 split
//        into a 15-bit head (k*ln2_hi exact) and a tail, then the Taylor
//        polynomial of e^r through r^7 (|r| <= ln2/2) in Horner form, then
//        scaling by 2^k.
//   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
//        s = f/(2+f), log m = f - (f^2/2 - s (f^2/2 + R(s^2))) with R the
//        atanh series 2s^2/3 + ... + 2s^10/11 (the reduction of fdlibm's
//        e_log.c, Sun Microsystems 1993).
// Both are within 2 ulp of the correctly rounded value over the normal range.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow2i(int k) {  // 2^k for k in [-126, 127]
    return __builtin_bit_cast(float, (uint32_t)(k + 127) << 23);
}

__device__ inline float det_expf(float x) {
#pragma clang fp contract(off)
    // k = rint(x / ln 2) clamped to a range where the int conversion is
    // defined (outside it the guards below decide); r = x - k ln2 with ln2
    // split hi + lo (k * hi is exact); Taylor through r^7 by explicit FMAs;
    // scaling by 2^k with one ldexp (correctly rounded, subnormals included).
    const float k = __builtin_amdgcn_fmed3f(__builtin_rintf(x * 1.44269502162933349609f), -200.0f, 200.0f);
    const float r = __builtin_fmaf(-k, 1.42860676533018704504e-06f, __builtin_fmaf(-k, 0.693145751953125f, x));
    float p = 1.98412698412698412698e-04f;
    p = __builtin_fmaf(p, r, 1.38888888888888888889e-03f);
    p = __builtin_fmaf(p, r, 8.33333333333333333333e-03f);
    p = __builtin_fmaf(p, r, 4.16666666666666666667e-02f);
    p = __builtin_fmaf(p, r, 1.66666666666666666667e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const float e = __builtin_ldexpf(p, (int)k);
    // guards as selects (no branch per call): NaN passes through, overflow
    // gives +inf, underflow below the smallest subnormal gives 0
    const float g = x > 88.72283935546875f ? __builtin_inff() : (x < -103.972084045410156f ? 0.0f : e);
    return x != x ? x : g;
}

// det_expf for x <= 0 or NaN (the softmax's z - max): the same value, with
// only the underflow guard (a NaN argument already yields NaN through the
// polynomial and ldexp, and no argument can overflow).
__device__ __forceinline__ float det_expf_nonpos(float x) {
#pragma clang fp contract(off)
    const float k = __builtin_amdgcn_fmed3f(__builtin_rintf(x * 1.44269502162933349609f), -200.0f, 200.0f);
    const float r = __builtin_fmaf(-k, 1.42860676533018704504e-06f, __builtin_fmaf(-k, 0.693145751953125f, x));
    float p = 1.98412698412698412698e-04f;
    p = __builtin_fmaf(p, r, 1.38888888888888888889e-03f);
    p = __builtin_fmaf(p, r, 8.33333333333333333333e-03f);
    p = __builtin_fmaf(p, r, 4.16666666666666666667e-02f);
    p = __builtin_fmaf(p, r, 1.66666666666666666667e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const float e = __builtin_ldexpf(p, (int)k);
    return x < -103.972084045410156f ? 0.0f : e;
}

// max(m, z) for the softmax's running maximum, which starts at 0: as signed
// integers the bit patterns of non-negative floats order like the floats and
// those of negative floats fall below 0, so the maximum over {0} and z is the
// integer maximum (fmaxf costs a canonicalisation on top in IEEE mode).
__device__ __forceinline__ float max_from_zero(float m, float z) {
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int, m), __builtin_bit_cast(int, z)));
}

__device__ inline float det_logf(float x) {
#pragma clang fp contract(off)
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e = 0;
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    if (b < 0x00800000u) {  // subnormal: scale into the normal range
        x = x * 16777216.0f;
        b = __builtin_bit_cast(uint32_t, x);
        e = -24;
    }
    e += (int)(b >> 23) - 127;
    float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f800000u);
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

// ---------------------------------------------------------------------------
// Encodings.  `in(d)` yields raw input d of the lane's row; `out(c, v)` stores
// encoded column c.  Formulas follow tiny-cuda-nn v1.6 (vendored in the
// reference under src/ext/tcnn).
// ---------------------------------------------------------------------------

// quartic_cdf, tcnn common_device.h:492-497: clamp(15/16 u (1 - 2u^2/3 + u^4/5)
// + 1/2, 0, 1), u = x * inv_radius; the polynomial by explicit fused
// multiply-adds (the same value to an ulp or two, fewer instructions).
__device__ __forceinline__ float quartic_cdf(float x, float inv_radius) {
#pragma clang fp contract(off)
    const float u = x * inv_radius;
    const float u2 = u * u;
    const float u4 = u2 * u2;
    const float poly = __builtin_fmaf(0.200000002980232238770f, u4, __builtin_fmaf(-0.666666686534881591797f, u2, 1.0f));
    return __builtin_amdgcn_fmed3f(__builtin_fmaf(0.9375f * u, poly, 0.5f), 0.0f, 1.0f);
}

// grid_index with the CoherentPrime hash, tcnn grid.h:111-128, :164-186.
// The dense stride is a uint32 as in the reference; when res^2 wraps to 0 the
// reference keeps the dense formula, and so does this.
__device__ __forceinline__ uint32_t grid_index3(int type, uint32_t hsize, uint32_t res, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t stride = 1, index = 0;
    if (stride <= hsize) { index += x * stride; stride *= res; }
    if (stride <= hsize) { index += y * stride; stride *= res; }
    if (stride <= hsize) { index += z * stride; stride *= res; }
    if (type == GRID_HASH && hsize < stride) index = x ^ (y * 2654435761u) ^ (z * 805459861u);
    return index % hsize;
}

// floor(v) as an int for any float: NaN -> 0, beyond the int range saturated
// (what v_cvt_i32_f32 does; written out so that no value is undefined in C++
// and the host restatement takes the same one).
__device__ __forceinline__ int floor_to_int(float v) {
    const float f = floorf(v);
    if (!(f == f)) return 0;
    return f >= 2147483648.0f ? 2147483647 : (f < -2147483648.0f ? (-2147483647 - 1) : (int)f);
}

// pos_fract, tcnn common_device.h:418-431 (identity or smoothstep), with the
// position p = in * scale + 0.5 taken exactly instead of rounded to fp32.
// Rounded, p carries up to half an ulp of error (2^-6 at scale 2^19), which
// is a percent of the fine levels' interpolation weights.  Here:
//   t = floor(fma(in, scale, 0.5))   (the fma rounds once; round-to-nearest
//                                     never lands below an integer <= p, so
//                                     t is floor(p) or, when p lies within
//                                     half an ulp below t, floor(p) + 1)
//   f = fma(in, scale, 0.5 - t)      (0.5 - t is exact below 2^24, so f is
//                                     p - t rounded once; f < 0 flags the
//                                     second case, which takes t - 1)
// so the cell is exactly floor(p) and the fraction p - cell within 2^-25,
// for |in * scale| < 2^23.
__device__ __forceinline__ void pos_fract(float in, float scale, int interp, float& frac, uint32_t& cell) {
#pragma clang fp contract(off)
    float t = floorf(__builtin_fmaf(in, scale, 0.5f));
    float p = __builtin_fmaf(in, scale, 0.5f - t);
    if (p < 0.0f) {
        t = t - 1.0f;
        p = __builtin_fmaf(in, scale, 0.5f - t);
    }
    cell = (uint32_t)floor_to_int(t);  // t is integral (or not finite)
    if (interp == INTERP_SMOOTHSTEP) p = p * p * (3.0f - 2.0f * p);
    frac = p;
}

template <int F>
struct FeatVec { float v[F]; };

template <int F>
__device__ __forceinline__ void load_feats(const _Float16* __restrict__ t, uint32_t idx, float (&o)[F]) {
    if constexpr (F == 1) {
        o[0] = (float)t[idx];
    } else if constexpr (F == 2) {
        h2 v = *(const h2*)(t + idx * 2);
        o[0] = (float)v[0]; o[1] = (float)v[1];
    } else if constexpr (F == 4) {
        h4 v = *(const h4*)(t + idx * 4);
#pragma unroll
        for (int f = 0; f < 4; ++f) o[f] = (float)v[f];
    } else {
        h8 v = *(const h8*)(t + idx * 8);
#pragma unroll
        for (int f = 0; f < 8; ++f) o[f] = (float)v[f];
    }
}

// One level of the grid for one row: the 8 corner indices and weights.
struct GridCorners {
    uint32_t idx[8];
    float w[8];
};

// n % hsize of one level: mask for powers of two, else Granlund & Montgomery's
// multiply-shift for an invariant divisor (exact for every uint32 n).
__device__ __forceinline__ uint32_t level_mod(const GridSpec& g, int level, uint32_t n, uint32_t hsize) {
    if (g.hashed[level] & 2u) return n & g.mod_mask[level];
    const uint32_t t = __umulhi(g.mod_mul[level], n);
    const uint32_t q = (t + ((n - t) >> 1)) >> g.mod_shift[level];
    return n - q * hsize;
}

// The 8 corner indices and weights of one level.  Same values as grid_index3
// on every corner, from the per-level closed form (hash or dense strides).
__device__ __forceinline__ void grid_corners(const GridSpec& g, int level, float px, float py, float pz, GridCorners& c) {
#pragma clang fp contract(off)
    const float scale = g.scale[level];
    const uint32_t hsize = g.offset[level + 1] - g.offset[level];
    float fx, fy, fz;
    uint32_t cx, cy, cz;
    pos_fract(px, scale, g.interp, fx, cx);
    pos_fract(py, scale, g.interp, fy, cy);
    pos_fract(pz, scale, g.interp, fz, cz);
    uint32_t ty[2], tz[2];
    const bool hashed = g.hashed[level] & 1u;
    if (hashed) {
        ty[0] = cy * 2654435761u; ty[1] = (cy + 1) * 2654435761u;
        tz[0] = cz * 805459861u;  tz[1] = (cz + 1) * 805459861u;
    } else {
        ty[0] = cy * g.s1[level]; ty[1] = (cy + 1) * g.s1[level];
        tz[0] = cz * g.s2[level]; tz[1] = (cz + 1) * g.s2[level];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float w = 1.0f;
        if (k & 1) { w *= fx; } else { w *= 1.0f - fx; }
        if (k & 2) { w *= fy; } else { w *= 1.0f - fy; }
        if (k & 4) { w *= fz; } else { w *= 1.0f - fz; }
        const uint32_t x = cx + (uint32_t)(k & 1);
        const uint32_t yz_h = ty[(k >> 1) & 1] ^ tz[k >> 2];
        const uint32_t yz_d = ty[(k >> 1) & 1] + tz[k >> 2];
        const uint32_t n = hashed ? (x ^ yz_h) : (x + yz_d);
        c.idx[k] = level_mod(g, level, n, hsize);
        c.w[k] = w;
    }
}

// Corners 2p and 2p + 1 (the x pair of the p-th (y, z) combination) of one
// level, by the same operations as grid_corners (identical indices and weights).
__device__ __forceinline__ void grid_corner_pair(const GridSpec& g, int level, float px, float py, float pz, int p,
                                                 uint32_t (&idx)[2], float (&w)[2]) {
#pragma clang fp contract(off)
    const float scale = g.scale[level];
    const uint32_t hsize = g.offset[level + 1] - g.offset[level];
    float fx, fy, fz;
    uint32_t cx, cy, cz;
    pos_fract(px, scale, g.interp, fx, cx);
    pos_fract(py, scale, g.interp, fy, cy);
    pos_fract(pz, scale, g.interp, fz, cz);
    const bool hashed = g.hashed[level] & 1u;
    const uint32_t yy = cy + (uint32_t)(p & 1), zz = cz + (uint32_t)(p >> 1);
    const uint32_t ty = hashed ? yy * 2654435761u : yy * g.s1[level];
    const uint32_t tz = hashed ? zz * 805459861u : zz * g.s2[level];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int k = 2 * p + j;
        float ww = 1.0f;
        if (k & 1) { ww *= fx; } else { ww *= 1.0f - fx; }
        if (k & 2) { ww *= fy; } else { ww *= 1.0f - fy; }
        if (k & 4) { ww *= fz; } else { ww *= 1.0f - fz; }
        const uint32_t x = cx + (uint32_t)j;
        const uint32_t n = hashed ? (x ^ (ty ^ tz)) : (x + (ty + tz));
        idx[j] = level_mod(g, level, n, hsize);
        w[j] = ww;
    }
}

// Grid level value for one row (N-linear / smoothstep / nearest).
template <int F>
__device__ __forceinline__ void grid_level(const GridSpec& g, const _Float16* __restrict__ table, int level,
                                           float px, float py, float pz, float (&res)[F]) {
#pragma clang fp contract(off)
    const _Float16* lt = table + (size_t)g.offset[level] * F;
    if (g.interp == INTERP_NEAREST) {
        GridCorners c;
        grid_corners(g, level, px, py, pz, c);  // corner 0 is the cell's own vertex
        load_feats<F>(lt, c.idx[0], res);
        return;
    }
    GridCorners c;
    grid_corners(g, level, px, py, pz, c);
    float v[8][F];
    if constexpr (F == 2) {
        // Fewer loads per lane: the x / x+1 corners usually sit in adjacent
        // entries (hash: x^h and x^h^1 for even x; dense: idx and idx+1), so
        // one 8-byte load serves both, and corner x+1 is loaded separately
        // only when it lies elsewhere.  On levels whose dense stride for z
        // wrapped to 0 the z+1 corners are the z corners.
        const uint32_t* lt32 = (const uint32_t*)lt;
        const bool zdup = !(g.hashed[level] & 1u) && g.s2[level] == 0u;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int k0 = 2 * p, k1 = 2 * p + 1;
            if (zdup && p >= 2) {
                v[k0][0] = v[k0 - 4][0]; v[k0][1] = v[k0 - 4][1];
                v[k1][0] = v[k1 - 4][0]; v[k1][1] = v[k1 - 4][1];
                continue;
            }
            // adjacent entries (either order) in one 8-byte load at the lower
            // index (4-byte aligned), else the aligned pair holding corner k0
            // and a separate load of corner k1
            const uint32_t i0 = c.idx[k0], i1 = c.idx[k1];
            const bool up = i1 == i0 + 1u, down = i1 + 1u == i0;
            uint2 pr;
            __builtin_memcpy(&pr, lt32 + (up ? i0 : (down ? i1 : (i0 & ~1u))), sizeof(pr));
            uint32_t e0, e1;
            if (up) {
                e0 = pr.x; e1 = pr.y;
            } else if (down) {
                e0 = pr.y; e1 = pr.x;
            } else {
                e0 = (i0 & 1u) ? pr.y : pr.x;
                e1 = lt32[i1];
            }
            const h2 h0 = __builtin_bit_cast(h2, e0), h1 = __builtin_bit_cast(h2, e1);
            v[k0][0] = (float)h0[0]; v[k0][1] = (float)h0[1];
            v[k1][0] = (float)h1[0]; v[k1][1] = (float)h1[1];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) load_feats<F>(lt, c.idx[k], v[k]);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) res[f] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int f = 0; f < F; ++f) res[f] += c.w[k] * v[k][f];
}

// One level of two grids of identical structure (the theta and phi nets) from
// their interleaved table: entry e is {theta f0 f1, phi f0 f1} in 8 bytes, so
// one line fetch serves both nets.  Same corner order and sums as grid_level<2>.
__device__ __forceinline__ void grid_level_pair(const GridSpec& g, const uint2* __restrict__ tp, int level,
                                                float px, float py, float pz, float (&rt)[2], float (&rp)[2]) {
#pragma clang fp contract(off)
    const uint2* lt = tp + g.offset[level];
    GridCorners c;
    grid_corners(g, level, px, py, pz, c);
    if (g.interp == INTERP_NEAREST) {
        const uint2 e = lt[c.idx[0]];
        const h2 a = __builtin_bit_cast(h2, e.x), b = __builtin_bit_cast(h2, e.y);
        rt[0] = (float)a[0]; rt[1] = (float)a[1];
        rp[0] = (float)b[0]; rp[1] = (float)b[1];
        return;
    }
    float vt[8][2], vp[8][2];
    const bool zdup = !(g.hashed[level] & 1u) && g.s2[level] == 0u;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int k0 = 2 * p, k1 = 2 * p + 1;
        if (zdup && p >= 2) {
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                vt[k0][f] = vt[k0 - 4][f]; vt[k1][f] = vt[k1 - 4][f];
                vp[k0][f] = vp[k0 - 4][f]; vp[k1][f] = vp[k1 - 4][f];
            }
            continue;
        }
        // adjacent entries (always on dense levels, and on hashed levels when
        // the x pair shares an aligned pair) come from one 16-byte load at
        // the lower index (8-byte aligned); otherwise the aligned pair holding
        // corner k0 and a separate load of corner k1
        const uint32_t i0 = c.idx[k0], i1 = c.idx[k1];
        const bool up = i1 == i0 + 1u, down = i1 + 1u == i0;
        const uint32_t base = up ? i0 : (down ? i1 : (i0 & ~1u));
        uint4 pr;
        __builtin_memcpy(&pr, lt + base, sizeof(pr));
        const uint2 lo = make_uint2(pr.x, pr.y), hi = make_uint2(pr.z, pr.w);
        uint2 e0, e1;
        if (up) {
            e0 = lo; e1 = hi;
        } else if (down) {
            e0 = hi; e1 = lo;
        } else {
            e0 = (i0 & 1u) ? hi : lo;
            e1 = lt[i1];
        }
        const h2 t0 = __builtin_bit_cast(h2, e0.x), p0 = __builtin_bit_cast(h2, e0.y);
        const h2 t1 = __builtin_bit_cast(h2, e1.x), p1 = __builtin_bit_cast(h2, e1.y);
        vt[k0][0] = (float)t0[0]; vt[k0][1] = (float)t0[1];
        vt[k1][0] = (float)t1[0]; vt[k1][1] = (float)t1[1];
        vp[k0][0] = (float)p0[0]; vp[k0][1] = (float)p0[1];
        vp[k1][0] = (float)p1[0]; vp[k1][1] = (float)p1[1];
    }
    rt[0] = rt[1] = rp[0] = rp[1] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            rt[f] += c.w[k] * vt[k][f];
            rp[f] += c.w[k] * vp[k][f];
        }
}

// Spherical harmonics up to degree 4 of (2x-1, 2y-1, 2z-1), tcnn
// spherical_harmonics.h:46-152 (the coefficients are the closed forms in its
// comments, e.g. 1/(2 sqrt(pi))).
template <typename Out>
__device__ __forceinline__ void sh_encode(int degree, float ix, float iy, float iz, Out out) {
#pragma clang fp contract(off)
    const float x = ix * 2.f - 1.f, y = iy * 2.f - 1.f, z = iz * 2.f - 1.f;
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    out(0, 0.28209479177387814f);
    if (degree <= 1) return;
    out(1, -0.48860251190291987f * y);
    out(2, 0.48860251190291987f * z);
    out(3, -0.48860251190291987f * x);
    if (degree <= 2) return;
    out(4, 1.0925484305920792f * xy);
    out(5, -1.0925484305920792f * yz);
    out(6, 0.94617469575755997f * z2 - 0.31539156525251999f);
    out(7, -1.0925484305920792f * xz);
    out(8, 0.54627421529603959f * x2 - 0.54627421529603959f * y2);
    if (degree <= 3) return;
    out(9, 0.59004358992664352f * y * (-3.0f * x2 + y2));
    out(10, 2.8906114426405538f * xy * z);
    out(11, 0.45704579946446572f * y * (1.0f - 5.0f * z2));
    out(12, 0.3731763325901154f * z * (5.0f * z2 - 3.0f));
    out(13, 0.45704579946446572f * x * (1.0f - 5.0f * z2));
    out(14, 1.4453057213202769f * z * (x2 - y2));
    out(15, 0.59004358992664352f * x * (-x2 + 3.0f * y2));
}

// Non-grid parts of encode_row.  NC > 0: the part's bin / frequency count
// as a compile-time constant (its loop unrolled, constant output offsets);
// NC == 0: the count from the part.  Both evaluate the same operations.
// (Unrolling the dimension loops as well measured slower: more registers.)

// kernel_one_blob_soa, tcnn oneblob.h:98-127: bin k gets S(b_{k+1}) - S(b_k),
// S(b) = sum over shifts s in {0,-1,+1} of Q(b - x + s).  Q is monotone and
// saturates (0 / 1) outside |t| < 1/n_bins, so for n_bins >= 2 only the shift
// nearest zero is unsaturated: S(b) = (1 + r) + Q(t - r), t = b - x,
// r = rint(t) clamped to [-1, 1] (t - r is exact, so Q sees the same argument
// as in the three-term sum).
template <int NC, typename In, typename Out>
__device__ __forceinline__ void oneblob_part(const EncPart& P, In& in, Out& out, int ob) {
#pragma clang fp contract(off)
    const int nb = NC > 0 ? NC : P.param;
    const float fnb = (float)nb;
    for (int d = 0; d < P.n_dims; ++d) {
        const float x = in(P.dim_begin + d);
        auto S = [&](float b) {
            const float t = b - x;
            if (nb < 2) return quartic_cdf(t, fnb) + quartic_cdf(t - 1.0f, fnb) + quartic_cdf(t + 1.0f, fnb);
            const float r = __builtin_amdgcn_fmed3f(__builtin_rintf(t), -1.0f, 1.0f);
            return (1.0f + r) + quartic_cdf(t - r, fnb);
        };
        float left = S(0.0f);
#pragma unroll
        for (int k = 0; k < (NC > 0 ? NC : nb); ++k) {
            const float right = S((float)(k + 1) / fnb);
            out(ob + d * nb + k, right - left);
            left = right;
        }
    }
}

// triangle_wave_encoding, tcnn triangle_wave.h:46-82
template <int NC, typename In, typename Out>
__device__ __forceinline__ void triwave_part(const EncPart& P, In& in, Out& out, int ob) {
#pragma clang fp contract(off)
    const int nf = NC > 0 ? NC : P.param;
    for (int d = 0; d < P.n_dims; ++d) {
        const float xin = in(P.dim_begin + d);
#pragma unroll
        for (int f = 0; f < (NC > 0 ? NC : nf); ++f) {
            const float x = ldexpf(xin, f - 1);
            const float val = x + (float)f * 0.25f;
            out(ob + d * nf + f, fabsf(val - floorf(val) - 0.5f) * 4.0f - 1.0f);
        }
    }
}

// frequency_encoding, tcnn frequency.h:46-94:
// sin(2^f pi x + (j%2) pi/2) = sin(2 pi r), r = frac(x 2^(f-1) + (j%2)/4);
// the reduction to one turn is exact, then v_sin_f32 (whose argument is in
// turns) evaluates the reduced angle.
template <int NC, typename In, typename Out>
__device__ __forceinline__ void frequency_part(const EncPart& P, In& in, Out& out, int ob) {
#pragma clang fp contract(off)
    const int nf = NC > 0 ? NC : P.param;
    for (int d = 0; d < P.n_dims; ++d) {
        const float xin = in(P.dim_begin + d);
#pragma unroll
        for (int j = 0; j < 2 * (NC > 0 ? NC : nf); ++j) {
            const float v = ldexpf(xin, (j >> 1) - 1) + (float)(j & 1) * 0.25f;
            const float fr = v - floorf(v);
            out(ob + d * 2 * nf + j, __builtin_amdgcn_sinf(fr));
        }
    }
}

// Whole composite encoding of one row.  In: callable in(d) -> float.
// Out: callable out(col, float).  Writes enc_width columns and zero-fills up to
// 32 * k_steps (the MFMA K padding, which has zero weights).
template <typename In, typename Out>
// gfeat (optional): this row's grid features already computed by the
// XCD-aware grid kernel, feature-major with column stride gstride; when given
// the grid part copies them instead of gathering.
// skip_grid: the grid part is neither gathered nor written (its features
// reach the MLP straight from the pre-pass buffer).
// pad_k: also write zeros from enc_width up to the K-step boundary (the MLP
// loader treats those columns as zero either way).
// part_first, part_step: encode only the nested parts part_first,
// part_first + part_step, ... (several lanes sharing one row's encoding).
__device__ __forceinline__ void encode_row(const NetSpec& ns, const _Float16* __restrict__ table, In in, Out out,
                                           const float* __restrict__ gfeat = nullptr, int64_t gstride = 0,
                                           bool skip_grid = false, bool pad_k = true, int part_first = 0,
                                           int part_step = 1) {
#pragma clang fp contract(off)
    for (int pi = part_first; pi < ns.n_parts; pi += part_step) {
        const EncPart& P = ns.parts[pi];
        const int ob = P.out_begin + (P.pad_front ? P.n_pad : 0);
        const int pb = P.pad_front ? P.out_begin : P.out_begin + P.n_out;
        for (int j = 0; j < P.n_pad; ++j) out(pb + j, 1.0f);
        switch (P.type) {
            case ENC_IDENTITY:
                for (int d = 0; d < P.n_dims; ++d) out(ob + d, in(P.dim_begin + d) * P.id_scale + P.id_offset);
                break;
            case ENC_ONEBLOB:
                // the common 4 bins with the bin loop unrolled
                if (P.param == 4) oneblob_part<4>(P, in, out, ob);
                else oneblob_part<0>(P, in, out, ob);
                break;
            case ENC_SH:
                sh_encode(P.param, in(P.dim_begin), in(P.dim_begin + 1), in(P.dim_begin + 2),
                          [&](int j, float v) { out(ob + j, v); });
                break;
            case ENC_TRIWAVE:
                if (P.param == 12) triwave_part<12>(P, in, out, ob);
                else triwave_part<0>(P, in, out, ob);
                break;
            case ENC_FREQUENCY:
                if (P.param == 12) frequency_part<12>(P, in, out, ob);
                else frequency_part<0>(P, in, out, ob);
                break;
            case ENC_GRID: {
                const GridSpec& g = ns.grid;
                const int F = g.n_feat;
                if (skip_grid) break;
                if (gfeat) {
                    for (int c = 0; c < g.n_levels * F; ++c) out(ob + c, gfeat[(size_t)c * gstride]);
                    break;
                }
                const float px = in(P.dim_begin), py = in(P.dim_begin + 1), pz = in(P.dim_begin + 2);
                for (int l = 0; l < g.n_levels; ++l) {
                    if (F == 2) {
                        float r[2]; grid_level<2>(g, table, l, px, py, pz, r);
                        out(ob + l * 2, r[0]); out(ob + l * 2 + 1, r[1]);
                    } else if (F == 4) {
                        float r[4]; grid_level<4>(g, table, l, px, py, pz, r);
#pragma unroll
                        for (int f = 0; f < 4; ++f) out(ob + l * 4 + f, r[f]);
                    } else if (F == 1) {
                        float r[1]; grid_level<1>(g, table, l, px, py, pz, r);
                        out(ob + l, r[0]);
                    } else {
                        float r[8]; grid_level<8>(g, table, l, px, py, pz, r);
#pragma unroll
                        for (int f = 0; f < 8; ++f) out(ob + l * 8 + f, r[f]);
                    }
                }
                break;
            }
        }
    }
    if (pad_k)
        for (int c = ns.enc_width; c < ns.k_steps * 32; ++c) out(c, 0.0f);
}

// ---------------------------------------------------------------------------
// DF head.  `P` is the lane's row of bins in LDS (stride 1).  All follow the
// reference's operation order with FMA contraction disabled.
// ---------------------------------------------------------------------------

// softmax2 + softmaxMultFactor_F, train.h:844-893 (fp32 storage): in-place on
// registers z[N] -> N * softmax(z), with the max initialised to 0.
// softmax2 / softmaxMultFactor_F (train.h:844-893): N exp(z_i - m) / sum_j
// exp(z_j - m), m = max(0, max z).  The reference normalises with a second
// exponential, exp(z_i - m - log sum); here each exponential is taken once and
// scaled by N / sum (the two agree to a few ulp; the oracle's FP32 mode does
// the same).  A sum that underflowed to 0 gives +inf bins, as the reference's
// log(0) does.
template <int N>
__device__ __forceinline__ void softmax_times_n(float (&z)[N]) {
#pragma clang fp contract(off)
    float mx = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) mx = max_from_zero(mx, z[i]);
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        z[i] = det_expf_nonpos(z[i] - mx);
        s += z[i];
    }
    const float scale = (float)N / s;
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = s > 0.0f ? z[i] * scale : __builtin_inff();
}

// Training variant: the reference stores the softmax in fp16 and then
// multiplies by (half)N in half precision (train.h:874-880).
template <int N>
__device__ __forceinline__ void softmax_times_n_half(float (&z)[N]) {
#pragma clang fp contract(off)
    float mx = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) mx = fmaxf(mx, z[i]);
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) s += det_expf(z[i] - mx);
    const float ls = det_logf(s);
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = (float)((_Float16)det_expf((z[i] - mx) - ls) * (_Float16)N);
}

struct Segment {
    int i;        // segment index (0..N for linear, 0..N for nearest)
    float c_lo;   // cdf at the segment start
    float c_hi;   // cdf at the segment end
};

// The searches' three selects per bin are kept in the bin's step: without
// this (empty) barrier the compiler splits them into three chains over all
// bins, keeps every bin's hit mask live in scalar registers for them (N mask
// pairs: spilled at 64 bins) and turns the index chain into float selects
// whose constants occupy vector registers.  No instruction is emitted.
__device__ __forceinline__ void select_barrier(Segment& s) {
    asm volatile("" : "+v"(s.i), "+v"(s.c_lo), "+v"(s.c_hi));
}

// search_sorted_cdf_linearinterpolation, train.h:361-397 — the sequential fp32
// CDF of the piecewise-linear density; branch-free, same sums in the same order.
template <int N>
__device__ __forceinline__ Segment search_linear(const float (&p)[N], float u) {
#pragma clang fp contract(off)
    // the first segment whose end exceeds u, tracked with selects (the same
    // result as the reference's early-exit loop, without a branch per bin)
    Segment s;
    float c = p[0] * 0.5f / (float)N;
    bool found = u < c;
    s.i = 0; s.c_lo = 0.0f; s.c_hi = c;
    float prev = c;
#pragma unroll
    for (int i = 1; i < N; ++i) {
        c = prev + (p[i - 1] + p[i]) / (float)(2 * N);
        const bool hit = (u < c) & !found;
        s.i = hit ? i : s.i;
        s.c_lo = hit ? prev : s.c_lo;
        s.c_hi = hit ? c : s.c_hi;
        found = found | hit;
        prev = c;
        select_barrier(s);
    }
    const float end = prev + p[N - 1] * 0.5f / (float)N;
    s.i = found ? s.i : N;
    s.c_lo = found ? s.c_lo : prev;
    s.c_hi = found ? s.c_hi : end;
    return s;
}

// search_sorted_cdf_nearest, train.h:345-359
template <int N>
__device__ __forceinline__ Segment search_nearest(const float (&p)[N], float u) {
#pragma clang fp contract(off)
    Segment s;
    float prev = 0.0f, c = 0.0f;
    bool found = false;
    s.i = N; s.c_lo = 0.0f; s.c_hi = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        c = prev + p[i] / (float)N;
        const bool hit = (u < c) & !found;
        s.i = hit ? i : s.i;
        s.c_lo = hit ? prev : s.c_lo;
        s.c_hi = hit ? c : s.c_hi;
        found = found | hit;
        prev = c;
        select_barrier(s);
    }
    s.c_lo = found ? s.c_lo : prev;
    s.c_hi = found ? s.c_hi : c;
    return s;
}

// Inverse CDF inside the found segment, sample_dir_guided_inputs_F,
// train.h:415-520.  `P` reads bin k (dynamic index).
template <int N, typename Bin>
__device__ __forceinline__ float sample_linear(const Segment& sg, float u, float u_extra, Bin P) {
#pragma clang fp contract(off)
    const float invN = 1.0f / (float)N;
    int vi = sg.i;
    int vim1 = vi - 1;
    float v = (float)vim1 + 0.5f;
    float ecm = 1.0f;
    float voff = 0.0f;
    if (vim1 < 0) {
        vim1 = vi; voff = 0.5f; ecm = 0.5f;
    } else if (vi > N - 1) {
        vi = vim1; ecm = 0.5f;
    }
    const float pa = P(vi), pb = P(vim1);
    const float m = (pa - pb) * (float)N;
    const float B = pb * (1.0f + v) - pa * v;
    v += voff;
    const float A = m * 0.5f;
    float x;
    if (fabsf(A) < DFG_EPS) {
        const float vd = u - sg.c_lo;
        const float cd = sg.c_hi - sg.c_lo;
        const float w = cd > DFG_EPS ? vd / cd : u_extra;
        x = (v + w * ecm) * invN;
    } else {
        // the branch is decided on the reference's global-coordinate
        // discriminant (train.h:494-497) ...
        const float C = -(m * 0.5f) * (v * v * invN * invN) - (B * v * invN) + sg.c_lo - u;
        const float delta = B * B - 4.0f * A * C;
        if (delta < DFG_EPS) {
            x = (v + u_extra * ecm) * invN;
        } else {
            // ... and the root taken in the segment's own coordinate t = x - x0,
            // x0 = v / N the left bin centre.  There the cdf is
            // c_lo + p0 t + A t^2 with p0 = P(vim1) (= m x0 + B), the
            // discriminant is p0^2 + 4 A d (= delta) with d = u - c_lo, and the
            // root the reference selects (the larger one for A >= 0, the
            // smaller for A < 0) is, for either sign, 2 d / (p0 + sqrt(p0^2 +
            // 4 A d)): free of the cancellation of the global form, whose
            // terms reach A x0^2 ~ N while the root moves x by ~1/N.
            const float d = u - sg.c_lo;
            const float disc = fmaxf(pb * pb + 4.0f * A * d, 0.0f);
            const float den = pb + __builtin_sqrtf(disc);
            const float t = den > 0.0f ? (2.0f * d) / den : 0.0f;
            x = v * invN + t;
        }
    }
    return fminf(x, 1.0f - DFG_EPS);
}

template <int N>
__device__ __forceinline__ float sample_nearest(const Segment& sg, float u, float u_extra) {
#pragma clang fp contract(off)
    const float vd = u - sg.c_lo;
    const float cd = sg.c_hi - sg.c_lo;
    const float w = cd > DFG_EPS ? vd / cd : u_extra;
    const float x = ((float)sg.i + w) / (float)N;
    return fminf(x, 1.0f - DFG_EPS);
}

// sample_pdf, train.h:654-681 (x is the square coordinate; idx = x * N).
// Total for every float x: on [0, 1] (and above it) the reference's indices;
// below the first bin centre's half-bin the reference's i1 runs negative and
// reads out of bounds, here it stays 0 (the constant end half-bin extends);
// NaN gives bin 0 with a NaN weight, i.e. a NaN pdf.
template <int N, typename Bin>
__device__ __forceinline__ float pdf_linear(float x, Bin P) {
#pragma clang fp contract(off)
    float idx = x * (float)N;
    idx -= 0.5f;
    int i0 = min(floor_to_int(idx), N - 1);
    int i1 = i0 + 1;
    float w = idx - (float)i0;
    if (i0 < 0) { i0 = N - 1; w = 1.0f; }
    if (i1 > N - 1) { i1 = 0; w = 0.0f; }
    i1 = max(i1, 0);
    return (1.0f - w) * P(i0) + w * P(i1);
}

template <int N, typename Bin>
__device__ __forceinline__ float pdf_nearest(float x, Bin P) {
    int i = min(floor_to_int(x * (float)N), N - 1);
    i = max(i, 0);  // the reference reads out of bounds for x < 0; clamp instead
    return P(i);
}

// add_theta_to_input_F, train.h:793-812: the phi net's theta input.
__device__ __forceinline__ float theta_input(float theta, int n_theta, int linear) {
#pragma clang fp contract(off)
    const float t = theta * (float)n_theta;
    const int fi = min(floor_to_int(t), n_theta - 1);
    if (linear) {
        const float w = t - (float)fi;
        return ((float)fi + w) / (float)n_theta;
    }
    return (float)fi / (float)n_theta;
}

// dL/dz_j for one net's logits (compute_gradients_specialized_linear_interpolation
// / _nearest, AR.h:50-104): multiplier * d pdf(x) / d z_j through the softmax
// Jacobian, where P(k) is the softmax probability of bin k (bins[k] / N).
// Per-row terms of combine_AR_pdf_and_dl_doutput (train.h:948-961) from the
// record t ({theta, phi, Li rgb, wiPdf}) and the nets' |pdf| at the sample:
// the guided pdf pdf() * pdfTransformInv() + eps as one fused multiply-add
// (the contraction CUDA applies to that expression), the gradient prefix and
// the row's loss.  ls_over_n: the loss scale over the batch size.
struct DfRowTerms {
    float prefix, loss;
};
__device__ __forceinline__ DfRowTerms df_row_terms(const float* t, float apt, float app, float ls_over_n) {
#pragma clang fp contract(off)
    const float Li = (t[2] + t[3] + t[4]) / 3.0f;
    const float wiPdf = t[5] + DFG_EPS;
    const float guided = __builtin_fmaf((apt * app) * DFG_INV_4PI, 1.0f / DFG_INV_4PI, DFG_EPS);
    DfRowTerms r;
    r.prefix = -Li / wiPdf / guided * ls_over_n;
    r.loss = -Li / wiPdf * det_logf(guided);
    return r;
}

// The same gradient one logit at a time: the constructor reads the (one or two)
// bins the pdf at x interpolates, operator() gives dL/dz_i from P(i).
template <int N>
struct DfLogitGrad {
    int i0, i1;
    float w, p0, p1;
    int linear;
    template <typename Prob>
    __device__ __forceinline__ DfLogitGrad(float x, int lin, Prob P) : linear(lin) {
#pragma clang fp contract(off)
        if (linear) {  // the indices of pdf_linear (total for every float x)
            const float idx = x * (float)N - 0.5f;
            i0 = min(floor_to_int(idx), N - 1);
            i1 = i0 + 1;
            w = idx - (float)i0;
            if (i0 < 0) { i0 = N - 1; w = 1.0f; }
            if (i1 > N - 1) { i1 = 0; w = 0.0f; }
            i1 = max(i1, 0);
            p0 = P(i0);
            p1 = P(i1);
        } else {
            i0 = min(floor_to_int(x * (float)N), N - 1);
            i0 = max(i0, 0);
            i1 = 0;
            w = 0.0f;
            p0 = P(i0);
            p1 = 0.0f;
        }
    }
    __device__ __forceinline__ float operator()(int i, float pi, float multiplier) const {
#pragma clang fp contract(off)
        if (linear) {
            const float j0 = i0 == i ? p0 * (1.0f - p0) : -p0 * pi;
            const float j1 = i1 == i ? p1 * (1.0f - p1) : -p1 * pi;
            const float j = (1.0f - w) * j0 + w * j1;
            return multiplier * j;
        }
        return multiplier * (i0 == i ? p0 * (1.0f - p0) : -p0 * pi);
    }
};

template <int N, typename Prob>
__device__ __forceinline__ void df_logit_grad(float x, float multiplier, int linear, Prob P, float (&g)[N]) {
    const DfLogitGrad<N> d(x, linear, P);
#pragma unroll
    for (int i = 0; i < N; ++i) g[i] = d(i, P(i), multiplier);
}

// ---------------------------------------------------------------------------
// MFMA helpers.  v_mfma_f32_16x16x32_f16: A 16x32 (lane l: row l%16, K slots
// 8 per lane), B 32x16 (lane l: column l%16, the same K slots), D 16x16 (lane
// l: column l%16, rows 4*(l/16)+r).  A = weights (rows = output neurons),
// B = activations (columns = rows of the batch), so each lane's D values are
// 4 neurons of one batch row and can feed the next layer without a transpose:
// the next layer's weight fragments are packed in the matching K order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ f4 mfma16(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ h8 ld_frag(const _Float16* __restrict__ pack, int t, int s, int ks, int lane) {
    return *(const h8*)(pack + ((size_t)((t * ks + s) * 64 + lane) << 3));
}

// Rounds a computed fp32 value to fp16 (to nearest even), as storing the
// finished fp32 result to half does in the reference.  The empty asm makes the
// value opaque: without it the compiler may fold the multiply that produced
// it into a mixed-precision multiply-add (v_fma_mixlo_f16) that rounds the
// exact product to fp16 once, which differs at and near fp16 rounding ties,
// and differs between kernels that inline the same encoder.
__device__ __forceinline__ _Float16 round_half(float v) {
    asm("" : "+v"(v));
    return (_Float16)v;
}

// fp16 conversion that saturates at +-65504 instead of overflowing to inf.
// Loss-scaled gradients of small batches can exceed the half range; the
// reference's plain conversion would turn them into inf and the backward into
// NaN.  (A NaN input comes out as -65504, fmaxf dropping the NaN: the loss
// heads pass their gradients through loss_grad first, so none arrives here
// from a loss.)
__device__ __forceinline__ _Float16 sat_h(float v) {
    return (_Float16)fminf(fmaxf(v, -65504.0f), 65504.0f);
}

// A loss head's gradient for one output: a value that is not finite (its
// row's loss term is not: e.g. a row whose logits all lie below -104, where
// the softmax with its maximum started at 0 (train.h:846) underflows to a zero
// sum, log 0, +inf bins and a NaN pdf, as the reference's does) contributes
// nothing.  The reference's NaN would reach every weight of the net through
// the backward and Adam; the row's loss value keeps the reference's NaN/inf,
// so the step's loss still shows it.
__device__ __forceinline__ float loss_grad(float g) {
    return fabsf(g) <= 3.402823466e+38f ? g : 0.0f;
}

// max(v, 0) as one integer max on the bit pattern: non-negative floats order
// like their bits, negative ones (and -0) have the sign bit and map to +0.
// (fmaxf costs a canonicalisation on top in IEEE mode.)
__device__ __forceinline__ float relu(float v) {
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int, v), 0));
}

// relu(C) of 4 M-tiles (64 neurons) -> the two K-step operands of the next layer.
__device__ __forceinline__ void relu_to_act(const f4 (&c)[4], h8 (&a)[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[s][j] = (_Float16)relu(c[2 * s + (j >> 2)][j & 3]);
}

__device__ __forceinline__ void c_to_act(const f4 (&c)[4], h8 (&a)[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) a[s][j] = sat_h(c[2 * s + (j >> 2)][j & 3]);
}

// Optional per-layer activation store (training): feature-major [64][ld] fp16.
// The first `steps` K-steps of layer 0 read straight from the grid pre-pass
// buffer (fp32, feature-major, column stride `stride`) instead of the LDS
// tile; rows past `last` are clamped to it (their results are discarded).
// The first K-step's grid features of a 64-row tile loaded ahead of the
// layer (lane: rows 16q + i16 of the four groups, features 8g..8g+7), so the
// loads overlap the work before it.
struct GridPre {
    float v[4][8];
};
struct GridDirect {
    const float* base;
    int64_t stride;
    int steps;
    int64_t row0;
    int64_t last;
    const GridPre* pre = nullptr;  // K-step 0 already loaded (all four groups in one pass)
};

// The first layer's weight fragments of K-step 0 (4 column tiles), loaded
// ahead of the layer by the caller so their latency overlaps earlier work.
struct W0Pre {
    h8 w[4];
};

__device__ __forceinline__ void w0_pre_load(const NetSpec& ns, const _Float16* __restrict__ fwd_pack, W0Pre& p) {
    const int lane = threadIdx.x & 63;
    const _Float16* pk = fwd_pack + ns.fwd_off[0];
#pragma unroll
    for (int t = 0; t < 4; ++t) p.w[t] = ld_frag(pk, t, 0, ns.k_steps, lane);
}

__device__ __forceinline__ void grid_pre_load(const GridDirect& gd, GridPre& p) {
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int64_t rr = gd.row0 + q * 16 + i16;
        if (rr > gd.last) rr = gd.last;
        const float* gsrc = gd.base + (size_t)(g * 8) * gd.stride + rr;
#pragma unroll
        for (int j = 0; j < 8; ++j) p.v[q][j] = gsrc[(size_t)j * gd.stride];
    }
}

struct ActStore {
    _Float16* x;        // encoded input, [k_steps*32][ld] (NULL: not stored)
    _Float16* act;      // hidden activations a_1..a_H, [H][64][ld]
    int64_t ld;         // leading dimension (padded batch)
    int64_t row0;       // first batch row of this tile
    _Float16* stage;    // LDS staging tile (64 x kTileS halves) for 16-byte stores of a
                        // whole 64-row tile, or NULL (element stores)
};

// The per-layer tiles of a 64-row block go through LDS ([feature][row], row
// stride kTileS halves), so that global memory sees 16-byte row runs of the
// feature-major [64][ld] activation and gradient arrays instead of one 2-byte
// access per element.
constexpr int kTileS = 72;

// act[0..63][row0 .. row0+63] -> tile (8 lanes per feature, 8 rows each).
__device__ __forceinline__ void tile_load(_Float16* tile, const _Float16* __restrict__ src, int64_t ld, int64_t row0,
                                          int lane) {
    const int c = lane & 7;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int n = (lane >> 3) + 8 * k;
        *(h8*)(tile + n * kTileS + 8 * c) = *(const h8*)(src + (size_t)n * ld + row0 + 8 * c);
    }
}

__device__ __forceinline__ void tile_store(const _Float16* tile, _Float16* __restrict__ dst, int64_t ld, int64_t row0,
                                           int lane) {
    const int c = lane & 7;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int n = (lane >> 3) + 8 * k;
        *(h8*)(dst + (size_t)n * ld + row0 + 8 * c) = *(const h8*)(tile + n * kTileS + 8 * c);
    }
}


// fp32 -> (hi, lo) fp16 pairs for two values at once, hi + lo == v to
// ~2^-23 relative (2^-25 absolute below the fp16 normal range): the MFMA then
// sees the activation at fp32-level precision through two products.
// hi = fp16(v) rounded to nearest (v_cvt_pk_f16_f32); v - hi is exact in
// fp32 (hi is within half an fp16 ulp of v), and v_fma_mix{lo,hi}_f16
// computes v * 1 + (-hi) with an fp16 operand and rounds that exact
// remainder to fp16 in one instruction, into the low or high half of `lo`
// (the mix instructions leave the other half alone; the second fills it).
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, h2{(_Float16)a, (_Float16)b});
    asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(lo) : "v"(a), "v"(hi));
    asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(lo) : "v"(b), "v"(hi));
}

// Eight values -> their (hi, lo) operands.
__device__ __forceinline__ void split8(const float (&v)[8], h8& hi, h8& lo) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) split_pair(v[2 * k], v[2 * k + 1], h[k], l[k]);
    hi = __builtin_bit_cast(h8, h);
    lo = __builtin_bit_cast(h8, l);
}

// relu(C) of 4 M-tiles -> (hi, lo) operands of the next layer.
__device__ __forceinline__ void relu_to_act2(const f4 (&c)[4], h8 (&hi)[2], h8 (&lo)[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = relu(c[2 * s + (j >> 2)][j & 3]);
        split8(v, hi[s], lo[s]);
    }
}

// One layer's activations (lane: rows q*16 + i16, neurons 16*(2s + j/4) +
// 4g + j%4) to the feature-major [64][ld] array: through the LDS staging tile
// as 16-byte row runs when the pass covers the whole 64-row tile, else element
// by element.
template <int QN>
__device__ __forceinline__ void store_act(const ActStore& st, _Float16* dst, const h8 (&act)[QN][2], int q0, int i16,
                                          int g, int lane) {
    if (QN == 4 && st.stage) {
#pragma unroll
        for (int q = 0; q < QN; ++q)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int n = 16 * (2 * s + (j >> 2)) + 4 * g + (j & 3);
                    st.stage[n * kTileS + q * 16 + i16] = act[q][s][j];
                }
        __syncthreads();
        tile_store(st.stage, dst, st.ld, st.row0, lane);
        __syncthreads();
        return;
    }
#pragma unroll
    for (int q = 0; q < QN; ++q)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = 16 * (2 * s + (j >> 2)) + 4 * g + (j & 3);
                dst[(size_t)n * st.ld + st.row0 + (q0 + q) * 16 + i16] = act[q][s][j];
            }
}

// Forward of a whole MLP for a 64-row tile.  xt: encoded input in LDS as fp32,
// row-major with stride xs floats (multiple of 4).  Writes the fp32 output of
// the last layer to `lo` (LDS, row-major stride ls floats).
// SPLIT: every activation (the encoded input included) enters the MFMA as an
// fp16 (hi, lo) pair, so the only rounding left is the fp32 accumulation and
// the fp16 weights themselves; otherwise activations are rounded to fp16
// (the reference's storage precision, used by the training forward).
template <int NT_OUT, bool SPLIT, int QN, typename TT>
__device__ __forceinline__ void mlp_forward_groups(const NetSpec& ns, const _Float16* __restrict__ fwd_pack,
                                                   const TT* xt, int xs, float* lo, int ls,
                                                   const ActStore* st, int q0, const GridDirect& gd,
                                                   const W0Pre* w0 = nullptr) {
    // an fp16 input tile has no low part to add
    constexpr bool SPLIT_IN = SPLIT && sizeof(TT) == 4;
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g = lane >> 4;
    f4 acc[QN][4];
    h8 act[QN][2];
    h8 actl[QN][2];
    // layer 0: B operand from the LDS tile, natural K order
    {
        const _Float16* pk = fwd_pack + ns.fwd_off[0];
#pragma unroll
        for (int q = 0; q < QN; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
        // K-step s + 1's weight fragments are loaded while step s computes
        // (step 0's may come prefetched by the caller)
        h8 wc[4];
        if (w0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) wc[t] = w0->w[t];
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) wc[t] = ld_frag(pk, t, 0, ns.k_steps, lane);
        }
        for (int s = 0; s < ns.k_steps; ++s) {
            h8 wn[4];
            const bool more = s + 1 < ns.k_steps;
            if (more) {
#pragma unroll
                for (int t = 0; t < 4; ++t) wn[t] = ld_frag(pk, t, s + 1, ns.k_steps, lane);
            }
            h8 b[QN], bl[QN];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                const TT* src = xt + ((q0 + q) * 16 + i16) * xs + (s - gd.steps) * 32 + g * 8;
                // columns from enc_width to the K-step boundary are zero and not in the tile
                const bool beyond = s * 32 + g * 8 >= ns.enc_width;
                if constexpr (sizeof(TT) == 2) {
                    // rows 16-byte aligned: one vector read; rows of an odd
                    // dword stride (per-lane rows without bank conflicts): 4 dwords
                    if (beyond) {
                        b[q] = h8{};
                    } else if ((xs & 7) == 0) {
                        b[q] = *(const h8*)src;
                    } else {
                        const uint32_t* s32 = (const uint32_t*)src;
                        uint4 u4 = make_uint4(s32[0], s32[1], s32[2], s32[3]);
                        b[q] = __builtin_bit_cast(h8, u4);
                    }
                } else {
                    float v[8];
                    if (s == 0 && gd.pre && QN == 4) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[j] = gd.pre->v[q][j];
                    } else if (s < gd.steps) {
                        int64_t rr = gd.row0 + (q0 + q) * 16 + i16;
                        if (rr > gd.last) rr = gd.last;
                        const float* gsrc = gd.base + (size_t)(s * 32 + g * 8) * gd.stride + rr;
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[j] = gsrc[(size_t)j * gd.stride];
                    } else if (beyond) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[j] = 0.0f;
                    } else if ((xs & 3) == 0) {
                        const float4 v0 = *(const float4*)src;
                        const float4 v1 = *(const float4*)(src + 4);
                        v[0] = v0.x; v[1] = v0.y; v[2] = v0.z; v[3] = v0.w;
                        v[4] = v1.x; v[5] = v1.y; v[6] = v1.z; v[7] = v1.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[j] = src[j];
                    }
                    if constexpr (SPLIT_IN) {
                        split8(v, b[q], bl[q]);
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; ++j) b[q][j] = (_Float16)v[j];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const h8 a = wc[t];
#pragma unroll
                for (int q = 0; q < QN; ++q) {
                    acc[q][t] = mfma16(a, b[q], acc[q][t]);
                    if (SPLIT_IN) acc[q][t] = mfma16(a, bl[q], acc[q][t]);
                }
            }
            if (more) {
#pragma unroll
                for (int t = 0; t < 4; ++t) wc[t] = wn[t];
            }
        }
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            if (SPLIT) relu_to_act2(acc[q], act[q], actl[q]);
            else relu_to_act(acc[q], act[q]);
        }
    }
    if (st) store_act<QN>(*st, st->act, act, q0, i16, g, lane);
    // hidden layers 1..H-1; the weight fragments of layer l + 1 are loaded
    // while layer l computes (they do not depend on the activations)
    h8 wf[4][2];
    auto load_layer = [&](int l, h8 (&w)[4][2]) {
        const _Float16* pk = fwd_pack + ns.fwd_off[l];
        const int nt = l == ns.n_hidden ? NT_OUT : 4;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < nt) {
                w[t][0] = ld_frag(pk, t, 0, 2, lane);
                w[t][1] = ld_frag(pk, t, 1, 2, lane);
            }
    };
    load_layer(1 < ns.n_hidden ? 1 : ns.n_hidden, wf);
    for (int l = 1; l < ns.n_hidden; ++l) {
        h8 wc[4][2];
#pragma unroll
        for (int t = 0; t < 4; ++t) { wc[t][0] = wf[t][0]; wc[t][1] = wf[t][1]; }
        load_layer(l + 1, wf);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const h8 a0 = wc[t][0];
            const h8 a1 = wc[t][1];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                f4 c = mfma16(a0, act[q][0], f4{0.f, 0.f, 0.f, 0.f});
                c = mfma16(a1, act[q][1], c);
                if (SPLIT) {
                    c = mfma16(a0, actl[q][0], c);
                    c = mfma16(a1, actl[q][1], c);
                }
                acc[q][t] = c;
            }
        }
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            if (SPLIT) relu_to_act2(acc[q], act[q], actl[q]);
            else relu_to_act(acc[q], act[q]);
        }
        if (st) store_act<QN>(*st, st->act + (size_t)l * 64 * st->ld, act, q0, i16, g, lane);
    }
    // output layer (its fragments are in wf)
    {
#pragma unroll
        for (int t = 0; t < NT_OUT; ++t) {
            const h8 a0 = wf[t][0];
            const h8 a1 = wf[t][1];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                f4 c = mfma16(a0, act[q][0], f4{0.f, 0.f, 0.f, 0.f});
                c = mfma16(a1, act[q][1], c);
                if (SPLIT) {
                    c = mfma16(a0, actl[q][0], c);
                    c = mfma16(a1, actl[q][1], c);
                }
                float* dst = lo + ((q0 + q) * 16 + i16) * ls + 16 * t + 4 * g;
#pragma unroll
                for (int r = 0; r < 4; ++r) dst[r] = c[r];
            }
        }
    }
}

// The 64-row tile runs through the MLP in passes of QN 16-row groups (fewer
// live accumulators per pass keeps the occupancy up; the weight fragments are
// re-read from L1 per pass).
template <int NT_OUT, bool SPLIT, int QN = (SPLIT ? 2 : 4), typename TT>
__device__ __forceinline__ void mlp_forward_tile(const NetSpec& ns, const _Float16* __restrict__ fwd_pack,
                                                 const TT* xt, int xs, float* lo, int ls,
                                                 const ActStore* st, const GridDirect& gd = GridDirect{nullptr, 0, 0, 0, 0},
                                                 const W0Pre* w0 = nullptr) {
    for (int q0 = 0; q0 < 4; q0 += QN)
        mlp_forward_groups<NT_OUT, SPLIT, QN>(ns, fwd_pack, xt, xs, lo, ls, st, q0, gd, q0 == 0 ? w0 : nullptr);
}

// ---------------------------------------------------------------------------
// XCD-aware grid encoding.  Workgroups are dispatched round-robin over the 8
// XCDs, so block b runs on XCD b % 8, and with n_groups dividing 8 the XCDs
// with b % n_groups == k serve level group k.  Blocks are ordered level-major
// in phases: in phase p, block b encodes level p * n_groups + b % n_groups for
// a chunk of rows, so each XCD works on one level (<= 4 MB of interleaved
// table) at a time and keeps it resident in its own L2, instead of streaming
// all 16 levels from the Infinity Cache.  One row per thread; positions from a
// compact [n][3] array; output feature-major fp32 [L*F][ld].
// ---------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ void grid_encode_block(const NetSpec& ns, const _Float16* __restrict__ table,
                                                  const float* __restrict__ pos, int64_t n, int64_t ld,
                                                  float* __restrict__ out, int n_groups, int chunks) {
    const GridSpec& g = ns.grid;
    const int b = blockIdx.x;
    const int phase = b / (n_groups * chunks);
    const int rem = b - phase * n_groups * chunks;
    const int level = phase * n_groups + rem % n_groups;
    const int64_t row = (int64_t)(rem / n_groups) * blockDim.x + threadIdx.x;
    if (row >= n || level >= g.n_levels) return;
    const float* p = pos + row * 3;
    float r[F];
    grid_level<F>(g, table, level, p[0], p[1], p[2], r);
#pragma unroll
    for (int f = 0; f < F; ++f) out[(size_t)(level * F + f) * ld + row] = r[f];
}

// The theta and phi pre-pass in one launch over their interleaved table
// (same XCD-aware level mapping as grid_encode_block).
__device__ __forceinline__ void grid_encode_pair_block(const NetSpec& ns, const uint2* __restrict__ tp,
                                                       const float* __restrict__ pos, int64_t n, int64_t ld,
                                                       float* __restrict__ out_th, float* __restrict__ out_ph,
                                                       int n_groups, int chunks, int b) {
    const GridSpec& g = ns.grid;
    const int phase = b / (n_groups * chunks);
    const int rem = b - phase * n_groups * chunks;
    const int level = phase * n_groups + rem % n_groups;
    const int64_t row = (int64_t)(rem / n_groups) * blockDim.x + threadIdx.x;
    if (row >= n || level >= g.n_levels) return;
    const float* p = pos + row * 3;
    float rt[2], rp[2];
    grid_level_pair(g, tp, level, p[0], p[1], p[2], rt, rp);
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        out_th[(size_t)(level * 2 + f) * ld + row] = rt[f];
        out_ph[(size_t)(level * 2 + f) * ld + row] = rp[f];
    }
}

}  // namespace dfg
