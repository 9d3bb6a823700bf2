// Host side of the guide: configuration -> network specs, parameter ownership
// (fp32 master, fp16 training copy, fp16 EMA inference copy, Adam state), the
// workspace, and every extern "C" entry point of include/dfg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ctx.h"
#include "json.h"

namespace dfg {


static std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// PCG32 (O'Neill 2014, XSH-RR output), used to seed the weights.
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    Pcg32(uint64_t seed, uint64_t seq) {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float next_float() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// ---------------------------------------------------------------------------
// Configuration -> NetSpec.  Composite width / padding rules follow
// tcnn encodings/composite.h:136-215 and encoding.h:85-87; grid level sizes
// follow grid.h:959-1025; the MLP layout follows fully_fused_mlp.cu:636-688.
// ---------------------------------------------------------------------------

static OptCfg parse_optimizer(const Json& j) {
    OptCfg o;
    const Json* adam = &j;
    std::string ot = lower(j.string_or("otype", "adam"));
    if (ot == "ema") {
        o.ema_decay = (float)j.number_or("decay", 0.99);
        if (!j.contains("nested")) throw Error(DFG_ERR_CONFIG, "EMA optimizer needs a nested optimizer");
        adam = &j.at("nested");
        ot = lower(adam->string_or("otype", "adam"));
    }
    if (ot != "adam") throw Error(DFG_ERR_CONFIG, "unsupported optimizer '" + ot + "' (Adam or EMA(Adam))");
    o.lr = (float)adam->number_or("learning_rate", 1e-3);
    o.beta1 = (float)adam->number_or("beta1", 0.9);
    o.beta2 = (float)adam->number_or("beta2", 0.999);
    o.eps = (float)adam->number_or("epsilon", 1e-8);
    o.l2 = (float)adam->number_or("l2_reg", 1e-8);
    if (adam->bool_or("adabound", false)) throw Error(DFG_ERR_CONFIG, "adabound is not supported");
    return o;
}

struct PartCfg {
    const Json* j;
    int type;
    int dims;
    int begin;
};

static int part_width(const Json& j, int type, int dims, int* param) {
    switch (type) {
        case ENC_IDENTITY: *param = 0; return dims;
        case ENC_ONEBLOB: {
            const int nb = (int)j.number_or("n_bins", 16);
            if (nb <= 0 || (nb & (nb - 1))) throw Error(DFG_ERR_CONFIG, "OneBlob n_bins must be a power of two");
            *param = nb;
            return dims * nb;
        }
        case ENC_SH: {
            const int deg = (int)j.number_or("degree", 4);
            if (dims != 3) throw Error(DFG_ERR_CONFIG, "SphericalHarmonics encodes exactly 3 dims");
            if (deg < 1 || deg > 4) throw Error(DFG_ERR_CONFIG, "SphericalHarmonics degree must be 1..4");
            *param = deg;
            return deg * deg;
        }
        case ENC_TRIWAVE: *param = (int)j.number_or("n_frequencies", 12); return dims * *param;
        case ENC_FREQUENCY: *param = (int)j.number_or("n_frequencies", 12); return dims * *param * 2;
        case ENC_GRID: {
            const int F = (int)j.number_or("n_features_per_level", 2);
            if (F != 1 && F != 2 && F != 4 && F != 8) throw Error(DFG_ERR_CONFIG, "grid n_features_per_level must be 1, 2, 4 or 8");
            if (dims != 3) throw Error(DFG_ERR_CONFIG, "grid encodings encode exactly 3 dims here");
            int L;
            if (j.contains("n_features") || j.contains("n_grid_features")) {
                const int nf = (int)(j.contains("n_features") ? j.at("n_features").num : j.at("n_grid_features").num);
                L = (nf + F - 1) / F;
            } else {
                L = (int)j.number_or("n_levels", 16);
            }
            if (L < 1 || L > DFG_MAX_LEVELS) throw Error(DFG_ERR_CONFIG, "grid n_levels out of range");
            *param = F;
            return L * F;
        }
    }
    throw Error(DFG_ERR_CONFIG, "bad encoding type");
}

static int enc_type(const std::string& ot) {
    const std::string s = lower(ot);
    if (s == "identity") return ENC_IDENTITY;
    if (s == "oneblob") return ENC_ONEBLOB;
    if (s == "sphericalharmonics") return ENC_SH;
    if (s == "trianglewave") return ENC_TRIWAVE;
    if (s == "frequency") return ENC_FREQUENCY;
    if (s == "grid" || s == "hashgrid" || s == "densegrid" || s == "tiledgrid") return ENC_GRID;
    throw Error(DFG_ERR_CONFIG, "unsupported encoding '" + ot + "'");
}

static NetSpec build_net(const Json& enc, const Json& net, int in_dims, int n_out) {
    NetSpec ns;
    std::memset(&ns, 0, sizeof(ns));
    ns.in_dims = in_dims;
    if (in_dims > DFG_MAX_IN) throw Error(DFG_ERR_CONFIG, "too many input dims");
    // nested list
    std::vector<const Json*> nested;
    if (lower(enc.string_or("otype", "OneBlob")) == "composite") {
        if (!enc.contains("nested") || !enc.at("nested").is_array()) throw Error(DFG_ERR_CONFIG, "Composite needs 'nested'");
        for (const auto& e : enc.at("nested").arr) nested.push_back(&e);
    } else {
        nested.push_back(&enc);
    }
    int specified = 0, unspecified = 0;
    for (auto* e : nested) {
        if (e->contains("n_dims_to_encode")) specified += (int)e->at("n_dims_to_encode").num;
        else ++unspecified;
    }
    if (unspecified > 1) throw Error(DFG_ERR_CONFIG, "only one nested encoding may omit n_dims_to_encode");
    if (specified > in_dims) throw Error(DFG_ERR_CONFIG, "nested encodings encode more dims than the input has");
    std::vector<PartCfg> parts;
    int offset = 0;
    for (auto* e : nested) {
        int dims = e->contains("n_dims_to_encode") ? (int)e->at("n_dims_to_encode").num : in_dims - specified;
        if (e->contains("dims_to_encode_begin")) offset = (int)e->at("dims_to_encode_begin").num;
        if (dims > 0) parts.push_back({e, enc_type(e->string_or("otype", "OneBlob")), dims, offset});
        offset += dims;
    }
    if (parts.empty() || (int)parts.size() > DFG_MAX_PARTS) throw Error(DFG_ERR_CONFIG, "1..6 nested encodings supported");
    ns.n_parts = (int)parts.size();
    // widths and padding (composite.h:188-199; last part padded to the network alignment 16)
    std::vector<int> width(parts.size()), param(parts.size()), align(parts.size());
    int n_grid = 0;
    for (size_t i = 0; i < parts.size(); ++i) {
        width[i] = part_width(*parts[i].j, parts[i].type, parts[i].dims, &param[i]);
        align[i] = parts[i].type == ENC_GRID ? param[i] : 1;
        if (parts[i].type == ENC_GRID) ++n_grid;
        if (parts[i].begin + parts[i].dims > in_dims) throw Error(DFG_ERR_CONFIG, "nested encoding reads past the input");
    }
    if (n_grid > 1) throw Error(DFG_ERR_CONFIG, "at most one grid encoding per network");
    int so_far = 0;
    int total_align = 16;
    for (int a : align) total_align = std::lcm(total_align, a);
    for (size_t i = 0; i < parts.size(); ++i) {
        int padded;
        if (i + 1 < parts.size()) {
            padded = (int)next_multiple((uint32_t)(so_far + width[i]), (uint32_t)align[i + 1]) - so_far;
        } else {
            padded = (int)next_multiple((uint32_t)(so_far + width[i]), (uint32_t)total_align) - so_far;
        }
        EncPart& P = ns.parts[i];
        P.type = parts[i].type;
        P.dim_begin = parts[i].begin;
        P.n_dims = parts[i].dims;
        P.out_begin = so_far;
        P.n_out = width[i];
        P.n_pad = padded - width[i];
        P.pad_front = P.type == ENC_SH ? 1 : 0;
        P.param = param[i];
        P.id_scale = (float)parts[i].j->number_or("scale", 1.0);
        P.id_offset = (float)parts[i].j->number_or("offset", 0.0);
        so_far += padded;
    }
    ns.enc_width = so_far;
    ns.k_steps = (so_far + 31) / 32;
    // a grid part that opens the encoding and fills whole K-steps can feed the
    // first MFMA layer straight from the pre-pass buffer
    ns.grid_ksteps = 0;
    if (ns.n_parts > 0 && ns.parts[0].type == ENC_GRID && ns.parts[0].n_pad == 0 && ns.parts[0].n_out % 32 == 0)
        ns.grid_ksteps = ns.parts[0].n_out / 32;
    if (ns.k_steps * 32 > DFG_MAX_ENC) throw Error(DFG_ERR_CONFIG, "encoded width above 128");
    // MLP
    const int width_n = (int)net.number_or("n_neurons", 64);
    if (width_n != DFG_WIDTH) throw Error(DFG_ERR_CONFIG, "only 64-neuron networks are supported");
    ns.n_hidden = (int)net.number_or("n_hidden_layers", 3);
    if (ns.n_hidden < 1 || ns.n_hidden > DFG_MAX_HIDDEN) throw Error(DFG_ERR_CONFIG, "n_hidden_layers must be 1..8");
    if (lower(net.string_or("activation", "ReLU")) != "relu") throw Error(DFG_ERR_CONFIG, "only ReLU activations");
    if (lower(net.string_or("output_activation", "None")) != "none") throw Error(DFG_ERR_CONFIG, "only a linear output");
    ns.n_out = n_out;
    ns.n_out_pad = (int)next_multiple((uint32_t)n_out, 16u);
    if (ns.n_out_pad > 64) throw Error(DFG_ERR_CONFIG, "at most 64 outputs");
    ns.n_layers = ns.n_hidden + 1;
    uint32_t off = 0;
    for (int l = 0; l <= ns.n_hidden; ++l) {
        ns.layer_off[l] = off;
        const uint32_t rows = l == ns.n_hidden ? (uint32_t)ns.n_out_pad : 64u;
        const uint32_t cols = l == 0 ? (uint32_t)ns.enc_width : 64u;
        off += rows * cols;
    }
    ns.layer_off[ns.n_hidden + 1] = off;
    ns.n_matrix_params = off;
    // grid
    for (int i = 0; i < ns.n_parts; ++i) {
        if (ns.parts[i].type != ENC_GRID) continue;
        const Json& j = *parts[i].j;
        GridSpec& g = ns.grid;
        g.present = 1;
        g.n_feat = ns.parts[i].param;
        g.n_levels = ns.parts[i].n_out / g.n_feat;
        const std::string ot = lower(j.string_or("otype", "Grid"));
        const std::string def_type = ot == "tiledgrid" ? "tiled" : (ot == "densegrid" ? "dense" : "hash");
        const std::string ty = lower(j.string_or("type", def_type));
        g.type = ty == "hash" ? GRID_HASH : (ty == "dense" ? GRID_DENSE : (ty == "tiled" ? GRID_TILED : -1));
        if (g.type < 0) throw Error(DFG_ERR_CONFIG, "grid type must be Hash, Dense or Tiled");
        if (lower(j.string_or("hash", "CoherentPrime")) != "coherentprime")
            throw Error(DFG_ERR_CONFIG, "only the CoherentPrime hash is supported");
        const std::string it = lower(j.string_or("interpolation", "Linear"));
        g.interp = it == "linear" ? INTERP_LINEAR : (it == "smoothstep" ? INTERP_SMOOTHSTEP : (it == "nearest" ? INTERP_NEAREST : -1));
        if (g.interp < 0) throw Error(DFG_ERR_CONFIG, "grid interpolation must be Linear, Smoothstep or Nearest");
        if (j.bool_or("stochastic_interpolation", false)) throw Error(DFG_ERR_CONFIG, "stochastic interpolation unsupported");
        g.base_res = (int)j.number_or("base_resolution", 16);
        g.per_level_scale = (float)j.number_or("per_level_scale", 2.0);
        const uint32_t log2T = (uint32_t)j.number_or("log2_hashmap_size", 19);
        const float l2s = std::log2(g.per_level_scale);
        uint32_t o = 0;
        for (int l = 0; l < g.n_levels; ++l) {
            const float scale = exp2f((float)l * l2s) * (float)g.base_res - 1.0f;
            const uint32_t res = (uint32_t)ceilf(scale) + 1;
            const uint32_t max_params = 0xffffffffu / 2;
            uint32_t pl = std::pow((float)res, 3.0f) > (float)max_params ? max_params : res * res * res;
            pl = next_multiple(pl, 8u);
            if (g.type == GRID_TILED) pl = std::min(pl, (uint32_t)(g.base_res * g.base_res * g.base_res));
            else if (g.type == GRID_HASH) pl = std::min(pl, 1u << log2T);
            g.scale[l] = scale;
            g.res[l] = res;
            g.offset[l] = o;
            o += pl;
        }
        g.offset[g.n_levels] = o;
        g.param_begin = ns.n_matrix_params;
        for (int l = 0; l < g.n_levels; ++l) {
            // closed form of grid_index's uint32 stride loop (grid.h:169-186)
            const uint32_t hsize = g.offset[l + 1] - g.offset[l];
            uint32_t stride = 1, s[3] = {0, 0, 0};
            for (int d = 0; d < 3 && stride <= hsize; ++d) {
                s[d] = stride;
                stride *= g.res[l];
            }
            g.s1[l] = s[1];
            g.s2[l] = s[2];
            const bool hashed = g.type == GRID_HASH && hsize < stride;
            const bool pow2 = (hsize & (hsize - 1)) == 0;
            g.hashed[l] = (hashed ? 1u : 0u) | (pow2 ? 2u : 0u);
            g.mod_mask[l] = hsize - 1;
            // Granlund & Montgomery (1994), unsigned division by an invariant d >= 2:
            // l = ceil(log2 d), m' = floor(2^32 (2^l - d) / d) + 1, q = (t + ((n - t) >> 1)) >> (l - 1)
            uint32_t lg = 0;
            while ((1ull << lg) < hsize) ++lg;
            if (!pow2) {
                g.mod_mul[l] = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << lg) - hsize)) / hsize + 1);
                g.mod_shift[l] = lg - 1;
            }
        }
    }
    ns.n_params = ns.n_matrix_params + (ns.grid.present ? ns.grid.offset[ns.grid.n_levels] * ns.grid.n_feat : 0);
    // fragment packs: forward then backward
    uint32_t p = 0;
    for (int l = 0; l <= ns.n_hidden; ++l) {
        ns.fwd_off[l] = p;
        const uint32_t mt = l == ns.n_hidden ? (uint32_t)ns.n_out_pad / 16 : 4u;
        const uint32_t ks = l == 0 ? (uint32_t)ns.k_steps : 2u;
        p += mt * ks * 512;
    }
    ns.fwd_off[ns.n_hidden + 1] = p;
    for (int l = 0; l <= ns.n_hidden; ++l) {
        ns.bwd_off[l] = p;
        const uint32_t mt = l == 0 ? (uint32_t)ns.enc_width / 16 : 4u;
        const uint32_t ks = l == ns.n_hidden ? (uint32_t)(ns.n_out_pad + 31) / 32 : 2u;
        p += mt * ks * 512;
    }
    ns.bwd_off[ns.n_hidden + 1] = p;
    ns.pack_size = p;
    return ns;
}

// ---------------------------------------------------------------------------
// Device state
// ---------------------------------------------------------------------------
}  // namespace dfg

using namespace dfg;

// the error of the calling thread's last failed dfg_create (no context to hold it)
static thread_local std::string g_create_error = "no error";

int dfg::fail(dfg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}



static void init_net_params(dfg_ctx* c, int id) {
    NetState& s = c->net[id];
    const NetSpec& ns = s.spec;
    std::vector<float> h(ns.n_params);
    Pcg32 rng(c->seed, (uint64_t)id + 1);
    for (int l = 0; l <= ns.n_hidden; ++l) {
        const uint32_t rows = l == ns.n_hidden ? (uint32_t)ns.n_out_pad : 64u;
        const uint32_t cols = l == 0 ? (uint32_t)ns.enc_width : 64u;
        const float bound = std::sqrt(6.0f / (float)(rows + cols));  // xavier uniform, gpu_matrix.h:291-306
        for (uint32_t i = 0; i < rows * cols; ++i) h[ns.layer_off[l] + i] = rng.next_float() * 2.0f * bound - bound;
    }
    for (size_t i = ns.n_matrix_params; i < ns.n_params; ++i)  // grid init U[-1e-4, 1e-4], grid.h:1364-1369
        h[i] = rng.next_float() * 2e-4f - 1e-4f;
    HIPX(hipMemcpy(s.master.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
}

static void sync_copies(dfg_ctx* c, int id, hipStream_t st) {
    NetState& s = c->net[id];
    launch_f32_to_f16(s.master.p, s.w16.p, s.spec.n_params, st);
    HIPX(hipMemcpyAsync(s.ema16.p, s.w16.p, s.spec.n_params * sizeof(_Float16), hipMemcpyDeviceToDevice, st));
    launch_pack(s.d_spec.p, s.spec, s.w16.p, s.pack_train.p, 3, st);
    launch_pack(s.d_spec.p, s.spec, s.ema16.p, s.pack_infer.p, 1, st);
    if (id < 2) refresh_pair_table(c, st);
    check_launch();
}

static void alloc_net(dfg_ctx* c, int id, const NetSpec& ns, const OptCfg& opt) {
    NetState& s = c->net[id];
    s.present = true;
    s.spec = ns;
    s.opt = opt;
    s.d_spec.ensure(1);
    HIPX(hipMemcpy(s.d_spec.p, &s.spec, sizeof(NetSpec), hipMemcpyHostToDevice));
    const size_t P = ns.n_params;
    s.master.ensure(P); s.m.ensure(P); s.v.ensure(P); s.grad.ensure(P);
    s.w16.ensure(P); s.ema16.ensure(P);
    // 16-bit step counters when 1 - beta^65535 is exactly 1.0f for both betas
    // (the device's exp2f(t log2 beta) <= 2^-26 there, with a bit to spare)
    const float l1 = (float)std::log2((double)opt.beta1), l2 = (float)std::log2((double)opt.beta2);
    s.short_steps = (float)kShortStepMax * l1 <= -26.0f && (float)kShortStepMax * l2 <= -26.0f;
    if (s.short_steps) s.steps16.ensure(P);
    else s.steps.ensure(P);
    s.pack_train.ensure(ns.pack_size);
    s.pack_infer.ensure(ns.fwd_off[ns.n_hidden + 1]);
    HIPX(hipMemset(s.m.p, 0, P * sizeof(float)));
    HIPX(hipMemset(s.v.p, 0, P * sizeof(float)));
    HIPX(hipMemset(s.grad.p, 0, P * sizeof(float)));
    if (s.short_steps) HIPX(hipMemset(s.steps16.p, 0, P * sizeof(uint16_t)));
    else HIPX(hipMemset(s.steps.p, 0, P * sizeof(uint32_t)));
    init_net_params(c, id);
    sync_copies(c, id, nullptr);
    HIPX(hipDeviceSynchronize());
}

extern "C" {

int dfg_create(const char* json_config, dfg_ctx** out_ctx) {
    if (!json_config || !out_ctx) return fail(nullptr, DFG_ERR_INVALID_ARG, "null argument");
    *out_ctx = nullptr;
    std::unique_ptr<dfg_ctx> c(new dfg_ctx());
    try {
        const Json root = parse_json(json_config);
        if (!root.is_object()) throw Error(DFG_ERR_CONFIG, "config must be a JSON object");
        Json empty;
        empty.kind = Json::Object;
        const Json& d = root.contains("dfg") ? root.at("dfg") : empty;
        c->n_theta = (int)d.number_or("n_bins_theta", 16);
        c->n_phi = (int)d.number_or("n_bins_phi", 32);
        const std::string interp = lower(d.string_or("interpolation", "linear"));
        if (interp != "linear" && interp != "nearest") throw Error(DFG_ERR_CONFIG, "interpolation must be linear or nearest");
        c->linear = interp == "linear";
        c->seed = (uint64_t)d.number_or("seed", 7272);
        c->D = (int)d.number_or("n_input_dims", 9);
        c->prepass_min_rows = (int64_t)d.number_or("grid_prepass_min_rows", 32768);
        c->query_chunks = (int)d.number_or("query_chunks", 1);
        c->prepass_blocks_per_cu = (int)d.number_or("prepass_blocks_per_cu", 0);
        c->pipeline_min_rows = (int64_t)d.number_or("pipeline_min_rows", 262144);
        if (c->query_chunks < 1 || c->query_chunks > 64) throw Error(DFG_ERR_CONFIG, "query_chunks must be in 1..64");
        if (c->prepass_blocks_per_cu < 0 || c->prepass_blocks_per_cu > 64)
            throw Error(DFG_ERR_CONFIG, "prepass_blocks_per_cu must be in 0..64");
        c->fused_training = d.number_or("fused_training", 1.0) != 0.0;
        c->keep_gradients = d.number_or("keep_gradients", 0.0) != 0.0;
        c->wgrad_chunk_rows = (int)d.number_or("wgrad_chunk_rows", 1024);
        if (c->wgrad_chunk_rows < 64 || c->wgrad_chunk_rows > 65536 || c->wgrad_chunk_rows % 64)
            throw Error(DFG_ERR_CONFIG, "wgrad_chunk_rows must be a multiple of 64 in 64..65536");
        {
            int dev = 0, cus = 0;
            HIPX(hipGetDevice(&dev));
            HIPX(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (cus > 0) c->n_cus = cus;
            int lds = 0;
            if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
                lds > 0)
                c->lds_per_cu = lds;
            c->device = dev;
            c->noalloc_scratch = noalloc_scratch_setup(dev);
        }
        const std::string gbw = lower(d.string_or("grid_backward", "bucketed"));
        if (gbw != "bucketed" && gbw != "atomic") throw Error(DFG_ERR_CONFIG, "grid_backward must be bucketed or atomic");
        c->grid_bwd_bucketed = gbw == "bucketed";
        auto okb = [](int b) { return b == 16 || b == 32 || b == 64; };
        if (!okb(c->n_theta) || !okb(c->n_phi)) throw Error(DFG_ERR_CONFIG, "bins must be 16, 32 or 64");
        const bool pair_ok = (c->n_theta == 16 && (c->n_phi == 16 || c->n_phi == 32)) ||
                             (c->n_theta == 32 && (c->n_phi == 32 || c->n_phi == 64)) ||
                             (c->n_theta == 64 && c->n_phi == 64);
        if (!pair_ok) throw Error(DFG_ERR_CONFIG, "unsupported (n_bins_theta, n_bins_phi) pair");
        if (c->D < 1 || c->D + 1 > DFG_MAX_IN) throw Error(DFG_ERR_CONFIG, "n_input_dims out of range");
        if (!root.contains("nn")) throw Error(DFG_ERR_CONFIG, "missing 'nn' section");
        const Json& nn = root.at("nn");
        // key lookup of learnedDistribution.cpp:84-90
        auto pick = [&](const char* a, const char* b) -> const Json& { return nn.contains(a) ? nn.at(a) : nn.at(b); };
        const NetSpec th = build_net(pick("theta_encoding", "encoding"), pick("theta_network", "network"), c->D, c->n_theta);
        const NetSpec ph = build_net(pick("phi_encoding", "encoding"), pick("phi_network", "network"), c->D + 1, c->n_phi);
        alloc_net(c.get(), DFG_NET_THETA, th, parse_optimizer(pick("theta_optimizer", "optimizer")));
        alloc_net(c.get(), DFG_NET_PHI, ph, parse_optimizer(pick("phi_optimizer", "optimizer")));
        {
            const GridSpec& a = c->net[0].spec.grid;
            const GridSpec& b = c->net[1].spec.grid;
            bool same = a.present && b.present && a.n_feat == 2 && b.n_feat == 2 && a.n_levels == b.n_levels &&
                        a.interp == b.interp && a.type == b.type &&
                        grid_pos_dim(c->net[0].spec) == grid_pos_dim(c->net[1].spec);
            for (int l = 0; same && l < a.n_levels; ++l)
                same = a.offset[l + 1] == b.offset[l + 1] && a.res[l] == b.res[l] && a.scale[l] == b.scale[l] &&
                       a.s1[l] == b.s1[l] && a.s2[l] == b.s2[l] && a.hashed[l] == b.hashed[l] &&
                       a.mod_mask[l] == b.mod_mask[l] && a.mod_mul[l] == b.mod_mul[l] &&
                       a.mod_shift[l] == b.mod_shift[l];
            c->pair_ok = same;
            refresh_pair_table(c.get(), nullptr);
            HIPX(hipDeviceSynchronize());
        }
        if (root.contains("nrc")) {
            const Json& nr = root.at("nrc");
            const NetSpec n = build_net(nr.at("encoding"), nr.at("network"), 15, 3);
            alloc_net(c.get(), DFG_NET_NRC, n, parse_optimizer(nr.at("optimizer")));
        }
    } catch (const Error& e) {
        return fail(nullptr, e.code, e.what());
    } catch (const std::exception& e) {
        return fail(nullptr, DFG_ERR_CONFIG, e.what());
    }
    *out_ctx = c.release();
    return DFG_OK;
}

int dfg_destroy(dfg_ctx* ctx) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    delete ctx;
    return DFG_OK;
}

const char* dfg_last_error(const dfg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int dfg_get_info(const dfg_ctx* ctx, dfg_info* info) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    if (!info) return DFG_ERR_INVALID_ARG;
    std::memset(info, 0, sizeof(*info));
    info->n_bins_theta = ctx->n_theta;
    info->n_bins_phi = ctx->n_phi;
    info->interpolation = ctx->linear;
    info->n_input_dims = ctx->D;
    info->has_nrc = ctx->net[DFG_NET_NRC].present ? 1 : 0;
    for (int i = 0; i < 3; ++i) {
        if (!ctx->net[i].present) continue;
        const NetSpec& s = ctx->net[i].spec;
        info->enc_width[i] = s.enc_width;
        info->n_hidden[i] = s.n_hidden;
        info->n_out_padded[i] = s.n_out_pad;
        info->n_params[i] = s.n_params;
        info->n_matrix_params[i] = s.n_matrix_params;
    }
    info->max_batch = (int64_t)1 << 30;
    return DFG_OK;
}

static NetState& need_net(dfg_ctx* c, int net) {
    if (net < 0 || net > 2) throw Error(DFG_ERR_INVALID_ARG, "unknown net id");
    if (!c->net[net].present) throw Error(DFG_ERR_INVALID_ARG, "net not configured");
    return c->net[net];
}

int dfg_get_params(dfg_ctx* ctx, int net, int which, float* host_out, size_t n) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    NetState& s = need_net(ctx, net);
    if (!host_out) throw Error(DFG_ERR_INVALID_ARG, "null output");
    if (n != s.spec.n_params) throw Error(DFG_ERR_BAD_SIZE, "parameter count mismatch");
    if (which == DFG_PARAMS_MASTER) {
        HIPX(hipMemcpy(host_out, s.master.p, n * sizeof(float), hipMemcpyDeviceToHost));
    } else if (which == DFG_PARAMS_TRAIN || which == DFG_PARAMS_INFERENCE) {
        ctx->tmp.ensure(n);
        launch_f16_to_f32(which == DFG_PARAMS_TRAIN ? s.w16.p : s.ema16.p, ctx->tmp.p, (int64_t)n, nullptr);
        check_launch();
        HIPX(hipMemcpy(host_out, ctx->tmp.p, n * sizeof(float), hipMemcpyDeviceToHost));
    } else {
        throw Error(DFG_ERR_INVALID_ARG, "unknown parameter view");
    }
    API_END(ctx)
}

int dfg_get_grad(dfg_ctx* ctx, int net, float* host_out, size_t n) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    NetState& s = need_net(ctx, net);
    if (!host_out) throw Error(DFG_ERR_INVALID_ARG, "null output");
    if (n != s.spec.n_params) throw Error(DFG_ERR_BAD_SIZE, "parameter count mismatch");
    if (!s.grad_kept)
        throw Error(DFG_ERR_INVALID_ARG,
                    "the grid part of this gradient went straight into the optimizer; create the context with "
                    "\"keep_gradients\": 1 to keep it");
    HIPX(hipMemcpy(host_out, s.grad.p, n * sizeof(float), hipMemcpyDeviceToHost));
    API_END(ctx)
}

int dfg_set_params(dfg_ctx* ctx, int net, const float* host_in, size_t n) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    NetState& s = need_net(ctx, net);
    if (!host_in) throw Error(DFG_ERR_INVALID_ARG, "null input");
    if (n != s.spec.n_params) throw Error(DFG_ERR_BAD_SIZE, "parameter count mismatch");
    HIPX(hipMemcpy(s.master.p, host_in, n * sizeof(float), hipMemcpyHostToDevice));
    sync_copies(ctx, net, nullptr);
    HIPX(hipDeviceSynchronize());
    API_END(ctx)
}

static const char* kFileNames[3] = {"AR_network_theta.bin", "AR_network_phi.bin", "NRC_network.bin"};

int dfg_save(dfg_ctx* ctx, const char* prefix) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    if (!prefix) return fail(ctx, DFG_ERR_INVALID_ARG, "null path");
    API_BEGIN
    for (int i = 0; i < 3; ++i) {
        if (!ctx->net[i].present) continue;
        std::vector<float> h(ctx->net[i].spec.n_params);
        HIPX(hipMemcpy(h.data(), ctx->net[i].master.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        const std::string path = std::string(prefix) + kFileNames[i];
        std::ofstream f(path, std::ios::binary);
        if (!f) throw Error(DFG_ERR_IO, "cannot open " + path);
        f.write((const char*)h.data(), (std::streamsize)(h.size() * sizeof(float)));
        if (!f) throw Error(DFG_ERR_IO, "write failed: " + path);
    }
    API_END(ctx)
}

int dfg_load(dfg_ctx* ctx, const char* prefix) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    if (!prefix) return fail(ctx, DFG_ERR_INVALID_ARG, "null path");
    API_BEGIN
    for (int i = 0; i < 3; ++i) {
        if (!ctx->net[i].present) continue;
        const std::string path = std::string(prefix) + kFileNames[i];
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) throw Error(DFG_ERR_IO, "cannot open " + path);
        const size_t bytes = (size_t)f.tellg();
        if (bytes != ctx->net[i].spec.n_params * sizeof(float))
            throw Error(DFG_ERR_BAD_SIZE, path + ": size does not match the network's parameter count");
        f.seekg(0);
        std::vector<float> h(ctx->net[i].spec.n_params);
        f.read((char*)h.data(), (std::streamsize)bytes);
        if (!f) throw Error(DFG_ERR_IO, "read failed: " + path);
        HIPX(hipMemcpy(ctx->net[i].master.p, h.data(), bytes, hipMemcpyHostToDevice));
        sync_copies(ctx, i, nullptr);
    }
    HIPX(hipDeviceSynchronize());
    API_END(ctx)
}

}  // extern "C"

// --- query family ---------------------------------------------------------------
void dfg::fill_lds_strides(const dfg_ctx* c, QueryArgs& a, bool with_nrc, bool with_df, int tiles_per_block) {
    // DF rows in LDS: the K-steps not read straight from a grid pre-pass buffer
    int ks = 0;
    for (int i = 0; i < 2; ++i) {
        const NetSpec& ns = c->net[i].spec;
        const bool direct = (i == 0 ? a.gfeat_th : a.gfeat_ph) != nullptr;
        ks = std::max(ks, ns.enc_width - (direct ? 32 * ns.grid_ksteps : 0));
    }
    // the tiles hold the encoded width only (columns up to the K-step boundary
    // are zero and never stored); +4 / +8 keeps 16-byte rows off one bank.
    // The output tile (stride ls floats) aliases the input tile and is written
    // 32 rows at a time while the next 32 input rows are still to be read, so
    // an input row must be at least as long as an output row.
    a.ls = std::max(std::max(c->n_theta, c->n_phi), 16) + 1;
    const int ls4 = (a.ls + 3) / 4 * 4;
    a.xs = with_df ? std::max((ks + 3) / 4 * 4 + 4, ls4) : 0;
    a.xs_nrc = with_nrc ? std::max((c->net[2].spec.enc_width + 7) / 8 * 8 + 8, 2 * ls4) : 0;
    a.lo_off = 0;
    a.ph_part0 = 0;
    // phi's nested parts after the grid that equal theta's (the same inputs,
    // widths and columns: SH and OneBlob of the same direction/normal) keep
    // theta's columns: the logits then go to the row's tail, past them, and
    // phi encodes only its remaining parts (TriangleWave of theta).  The
    // longer row is taken only while the CU still holds at least 8 tiles
    // (2 waves per SIMD) of the kernel's workgroups: with 32 bins (row 69
    // floats) the 4-tile query keeps 2 workgroups per CU and gains; with 64
    // bins the row would grow to 101 floats and drop the query to 1 workgroup
    // per CU (1M queries, 64/64 bins: 0.85 ms with the shared columns, 0.70
    // without; 6-bounce frame 4.50 -> 4.22 ms).
    auto tile_bytes = [&](int xs, int ls) {
        int64_t b = (int64_t)64 * (xs | 1) * 4;
        b = std::max<int64_t>(b, (int64_t)64 * (ls | 1) * 4);
        b = std::max<int64_t>(b, (int64_t)64 * a.xs_nrc * 2);
        return (b + 15) / 16 * 16;
    };
    auto tiles_per_cu = [&](int xs, int ls) {
        const int tpb = std::max(1, tiles_per_block);
        return (c->lds_per_cu / (tpb * tile_bytes(xs, ls))) * tpb;
    };
    const NetSpec& th = c->net[0].spec;
    const NetSpec& ph = c->net[1].spec;
    if (with_df && a.gfeat_th && a.gfeat_ph && th.grid_ksteps > 0 && th.grid_ksteps == ph.grid_ksteps &&
        th.n_parts >= 2 && ph.n_parts > th.n_parts) {
        bool same = true;
        for (int i = 1; i < th.n_parts && same; ++i) {
            const EncPart& p = th.parts[i];
            const EncPart& q = ph.parts[i];
            same = p.type == q.type && p.dim_begin == q.dim_begin && p.n_dims == q.n_dims &&
                   p.out_begin == q.out_begin && p.n_out == q.n_out && p.pad_front == q.pad_front &&
                   p.param == q.param && p.id_scale == q.id_scale && p.id_offset == q.id_offset &&
                   p.type != ENC_GRID && (i + 1 == th.n_parts || p.n_pad == q.n_pad) && (p.pad_front == 0 || p.n_pad == 0);
        }
        if (same) {
            const EncPart& last = th.parts[th.n_parts - 1];
            const int end = last.out_begin + last.n_out - 32 * th.grid_ksteps;  // shared columns of the row
            const int lo_off = (end + 3) / 4 * 4;
            const int xs = std::max(a.xs, lo_off + ls4 + 4);
            if (tiles_per_cu(xs, xs) >= 8) {
                a.lo_off = lo_off;
                a.xs = xs;
                a.ls = a.xs;  // the logits share the input row's stride
                a.ph_part0 = th.n_parts;
            }
        }
    }
    // one row per lane: the lanes' same-column accesses (encoder stores, the
    // heads' bin reads and writes) hit 64 different banks when the row stride
    // in dwords is odd (the first layer then reads its operands as dwords)
    if (a.xs) {
        a.xs = a.xs | 1;
        if (a.lo_off) a.ls = a.xs;
        else a.ls = std::max(a.ls, 1) | 1;
    }
}

QueryArgs dfg::base_args(dfg_ctx* c) {
    QueryArgs a;
    std::memset(&a, 0, sizeof(a));
    a.th = c->net[0].d_spec.p;
    a.ph = c->net[1].d_spec.p;
    a.pack_th = c->net[0].pack_infer.p;
    a.pack_ph = c->net[1].pack_infer.p;
    a.tab_th = c->net[0].tab_infer();
    a.tab_ph = c->net[1].tab_infer();
    if (c->net[2].present) {
        a.nrc = c->net[2].d_spec.p;
        a.pack_nrc = c->net[2].pack_infer.p;
        a.tab_nrc = c->net[2].tab_infer();
    }
    a.linear = c->linear;
    return a;
}

// XCD-aware grid pre-pass of both DF nets for n rows (when worth a launch).
// Compact copy of the grid positions (dims pos_dim..pos_dim+2) for the pre-pass.
// buf: the context's query positions or the training step's own.
static const float* compact_pos(dfg::DevBuf<float>& buf, const float* feat, int stride, int pos_dim, int64_t n,
                                hipStream_t st) {
    buf.ensure((size_t)n * 3);
    launch_extract_pos(feat, stride, pos_dim, n, buf.p, st);
    return buf.p;
}

// Rebuilds both interleaved tables from the nets' own (after parameters were
// set or loaded; the training step's Adam keeps them current itself).
void dfg::refresh_pair_table(dfg_ctx* c, hipStream_t st) {
    if (!c->pair_ok) return;
    const NetSpec& ns = c->net[0].spec;
    const int64_t entries = ns.grid.offset[ns.grid.n_levels];
    c->tab_pair.ensure((size_t)entries);
    c->tab_pair_train.ensure((size_t)entries);
    launch_interleave_tables(c->net[0].tab_infer(), c->net[1].tab_infer(), c->tab_pair.p, entries, st);
    launch_interleave_tables(c->net[0].tab_train(), c->net[1].tab_train(), c->tab_pair_train.p, entries, st);
}

dfg::PrepassSizes dfg::prepass_sizes(const dfg_ctx* c, int64_t n) {
    PrepassSizes z;
    z.ld = round_up(n, 64);
    z.pos = (size_t)n * 3;
    for (int i = 0; i < 2; ++i) {
        const NetSpec& ns = c->net[i].spec;
        z.gfeat[i] = ns.grid.present ? (size_t)ns.grid.n_levels * (c->pair_ok ? 2 : ns.grid.n_feat) * z.ld : 0;
    }
    return z;
}

void dfg::grid_prepass_into(dfg_ctx* c, QueryArgs& a, int64_t n, const float* feat, int stride, hipStream_t st,
                            float* pos_buf, float* gfeat_th, float* gfeat_ph) {
    const int64_t ld = round_up(n, 64);
    float* gf[2] = {gfeat_th, gfeat_ph};
    if (c->pair_ok) {
        launch_extract_pos(feat, stride, grid_pos_dim(c->net[0].spec), n, pos_buf, st);
        launch_grid_encode_pair(c->net[0].d_spec.p, c->net[0].spec, c->tab_pair.p, pos_buf, n, ld, gf[0], gf[1], st);
        a.gfeat_th = gf[0];
        a.gfeat_ph = gf[1];
        a.gld = ld;
        return;
    }
    int pos_dim = -1;
    for (int i = 0; i < 2; ++i) {
        NetState& s = c->net[i];
        if (!s.spec.grid.present) continue;
        const int pd = grid_pos_dim(s.spec);
        if (pd != pos_dim) {  // stream order: the previous net's pre-pass has read the buffer
            launch_extract_pos(feat, stride, pd, n, pos_buf, st);
            pos_dim = pd;
        }
        launch_grid_encode(s.d_spec.p, s.spec, s.tab_infer(), pos_buf, n, ld, gf[i], st);
        (i == 0 ? a.gfeat_th : a.gfeat_ph) = gf[i];
    }
    a.gld = ld;
}

void dfg::grid_prepass(dfg_ctx* c, QueryArgs& a, int64_t n, const float* feat, int stride, hipStream_t st) {
    if (n < c->prepass_min_rows) return;
    const PrepassSizes z = prepass_sizes(c, n);
    c->pos.ensure(z.pos);
    for (int i = 0; i < 2; ++i)
        if (z.gfeat[i]) c->net[i].gfeat.ensure(z.gfeat[i]);
    grid_prepass_into(c, a, n, feat, stride, st, c->pos.p, c->net[0].gfeat.p, c->net[1].gfeat.p);
}

// Chunked query of a large batch: the paired grid pre-pass of every chunk is
// queued on a side stream, and the query tiles of chunk k (on the caller's
// stream) wait only for the pre-pass of chunk k.  The pre-pass is bound by L2
// requests (random table lines) and the tiles by VALU work, so the pre-pass of
// the next chunks runs on the CUs beside the tiles of the current one.  The
// per-row results are those of the unchunked query.  Returns false (nothing
// launched) when the batch is not chunked.
static bool query_chunked(dfg_ctx* c, const QueryArgs& base, int64_t n, hipStream_t st) {
    if (!c->pair_ok || c->query_chunks < 2 || n < c->pipeline_min_rows || n < c->prepass_min_rows) return false;
    if (!c->side) {
        HIPX(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPX(hipEventCreateWithFlags(&c->pipe_start, hipEventDisableTiming));
    }
    while ((int)c->pipe_ev.size() < c->query_chunks) {
        hipEvent_t e;
        HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->pipe_ev.push_back(e);
    }
    const int64_t ld = round_up(n, 64);
    const size_t cols = (size_t)c->net[0].spec.grid.n_levels * 2;
    c->net[0].gfeat.ensure(cols * ld);
    c->net[1].gfeat.ensure(cols * ld);
    float* gth = c->net[0].gfeat.p;
    float* gph = c->net[1].gfeat.p;
    const float* pos = compact_pos(c->pos, base.feat, base.stride, grid_pos_dim(c->net[0].spec), n, st);
    // the side stream starts after everything queued so far on the caller's
    // stream (an earlier query may still read the feature buffers)
    HIPX(hipEventRecord(c->pipe_start, st));
    HIPX(hipStreamWaitEvent(c->side, c->pipe_start, 0));
    const int64_t m = round_up((n + c->query_chunks - 1) / c->query_chunks, 256);
    const int max_blocks = c->prepass_blocks_per_cu * c->n_cus;
    int k = 0;
    for (int64_t r0 = 0; r0 < n; r0 += m, ++k) {
        const int64_t rows = std::min(m, n - r0);
        launch_grid_encode_pair(c->net[0].d_spec.p, c->net[0].spec, c->tab_pair.p, pos + 3 * r0, rows, ld, gth + r0,
                                gph + r0, c->side, nullptr, max_blocks);
        HIPX(hipEventRecord(c->pipe_ev[k], c->side));
    }
    k = 0;
    for (int64_t r0 = 0; r0 < n; r0 += m, ++k) {
        const int64_t rows = std::min(m, n - r0);
        HIPX(hipStreamWaitEvent(st, c->pipe_ev[k], 0));
        QueryArgs a = base;
        a.feat = base.feat + r0 * base.stride;
        a.n = rows;
        if (base.u) a.u = base.u + 4 * r0;
        if (base.guided) a.guided = base.guided + r0;
        if (base.wi) a.wi = base.wi + 4 * r0;
        if (base.rgb) a.rgb = base.rgb + 3 * r0;
        if (base.wi2) a.wi2 = base.wi2 + 2 * r0;
        if (base.pdf_out) a.pdf_out = base.pdf_out + r0;
        a.gfeat_th = gth + r0;
        a.gfeat_ph = gph + r0;
        a.gld = ld;
        fill_lds_strides(c, a, base.rgb != nullptr, true);
        launch_query(a, c->n_theta, c->n_phi, st);
    }
    return true;
}

void dfg::check_rows(const dfg_ctx* c, int64_t n, const float* feat, int stride) {
    if (n < 0 || n > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "n out of range");
    if (n > 0 && !feat) throw Error(DFG_ERR_INVALID_ARG, "null features");
    if (stride < c->D) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below n_input_dims");
}

extern "C" {

int dfg_query(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride, const float* u,
              const uint8_t* guided, float* wi, float* nrc_rgb) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, n, feat, feat_stride);
    if (n > 0 && (!u || !wi)) throw Error(DFG_ERR_INVALID_ARG, "null u or wi");
    if (nrc_rgb && !ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (nrc_rgb && feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "NRC output needs feat_stride >= 15");
    if (n == 0) return DFG_OK;
    QueryArgs a = base_args(ctx);
    a.feat = feat; a.stride = feat_stride; a.n = n; a.u = u; a.guided = guided; a.wi = wi; a.rgb = nrc_rgb;
    if (query_chunked(ctx, a, n, (hipStream_t)stream)) {
        check_launch();
        return DFG_OK;
    }
    grid_prepass(ctx, a, n, feat, feat_stride, (hipStream_t)stream);
    fill_lds_strides(ctx, a, nrc_rgb != nullptr, true);
    launch_query(a, ctx->n_theta, ctx->n_phi, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

int dfg_pdf(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride, const float* wi2, float* pdf) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, n, feat, feat_stride);
    if (n > 0 && (!wi2 || !pdf)) throw Error(DFG_ERR_INVALID_ARG, "null wi2 or pdf");
    if (n == 0) return DFG_OK;
    QueryArgs a = base_args(ctx);
    a.feat = feat; a.stride = feat_stride; a.n = n; a.wi2 = wi2; a.pdf_out = pdf;
    if (query_chunked(ctx, a, n, (hipStream_t)stream)) {
        check_launch();
        return DFG_OK;
    }
    grid_prepass(ctx, a, n, feat, feat_stride, (hipStream_t)stream);
    fill_lds_strides(ctx, a, false, true);
    launch_query(a, ctx->n_theta, ctx->n_phi, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

int dfg_logits(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride, const float* theta_in,
               float* logits_theta, float* logits_phi) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, n, feat, feat_stride);
    if (n > 0 && (!theta_in || !logits_theta || !logits_phi)) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return DFG_OK;
    QueryArgs a = base_args(ctx);
    a.feat = feat; a.stride = feat_stride; a.n = n; a.theta_in = theta_in;
    a.logits_th = logits_theta; a.logits_ph = logits_phi;
    grid_prepass(ctx, a, n, feat, feat_stride, (hipStream_t)stream);
    fill_lds_strides(ctx, a, false, true);
    launch_query(a, ctx->n_theta, ctx->n_phi, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

int dfg_nrc_infer(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride, float* rgb) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (n < 0 || n > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "n out of range");
    if (feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below 15");
    if (n > 0 && (!feat || !rgb)) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (n == 0) return DFG_OK;
    QueryArgs a = base_args(ctx);
    a.th = nullptr;
    a.feat = feat; a.stride = feat_stride; a.n = n; a.rgb = rgb;
    fill_lds_strides(ctx, a, true, false);
    launch_query(a, ctx->n_theta, ctx->n_phi, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

int dfg_query_plain(dfg_ctx* ctx, void* stream, int64_t n, const float* feat, int feat_stride, const float* u,
                    const uint8_t* guided, float* wi, float* nrc_rgb) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, n, feat, feat_stride);
    if (n > 0 && (!u || !wi)) throw Error(DFG_ERR_INVALID_ARG, "null u or wi");
    if (nrc_rgb && !ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (nrc_rgb && feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "NRC output needs feat_stride >= 15");
    if (n == 0) return DFG_OK;
    hipStream_t st = (hipStream_t)stream;
    // the reference synchronises the device at the start of every query stage
    HIPX(hipStreamSynchronize(st));
    const int64_t ld = round_up(n, 64);
    int kmax = std::max(ctx->net[0].spec.enc_width, ctx->net[1].spec.enc_width);
    if (nrc_rgb) kmax = std::max(kmax, ctx->net[2].spec.enc_width);
    ctx->p_enc.ensure((size_t)kmax * ld);
    ctx->p_a.ensure((size_t)64 * ld);
    ctx->p_b.ensure((size_t)64 * ld);
    ctx->p_theta.ensure((size_t)ld);
    auto run_mlp = [&](NetState& s, const float* extra) -> const float* {
        const NetSpec& ns = s.spec;
        launch_plain_encode(s.d_spec.p, s.tab_infer(), feat, feat_stride, extra, n, ld, ctx->p_enc.p, st);
        const float* X = ctx->p_enc.p;
        float* bufs[2] = {ctx->p_a.p, ctx->p_b.p};
        int cur = 0;
        for (int l = 0; l <= ns.n_hidden; ++l) {
            const int M = l == ns.n_hidden ? ns.n_out_pad : 64;
            const int K = l == 0 ? ns.enc_width : 64;
            launch_plain_layer(s.ema16.p + ns.layer_off[l], M, K, K, X, bufs[cur], l < ns.n_hidden, n, ld, st);
            X = bufs[cur];
            cur ^= 1;
        }
        return X;
    };
    const float* lt = run_mlp(ctx->net[0], nullptr);
    launch_plain_head(0, ctx->n_theta, ctx->n_theta, ctx->linear, lt, ld, n, u, guided, wi, ctx->p_theta.p, st);
    const float* lp = run_mlp(ctx->net[1], ctx->p_theta.p);
    launch_plain_head(1, ctx->n_phi, ctx->n_theta, ctx->linear, lp, ld, n, u, guided, wi, ctx->p_theta.p, st);
    if (nrc_rgb) {
        const float* o = run_mlp(ctx->net[2], nullptr);
        launch_plain_rgb(o, ld, n, nrc_rgb, st);
    }
    check_launch();
    API_END(ctx)
}

// --- training ----------------------------------------------------------------------
static void ensure_train_ws(NetState& s, int64_t ld, int n_chunks) {
    const NetSpec& ns = s.spec;
    s.x.ensure((size_t)ns.k_steps * 32 * ld);
    s.act.ensure((size_t)ns.n_hidden * 64 * ld);
    s.g.ensure((size_t)ns.n_hidden * 64 * ld);
    s.out.ensure((size_t)ns.n_out_pad * ld);
    s.gout_rm.ensure((size_t)ns.n_out_pad * ld);
    s.gout_fm.ensure((size_t)ns.n_out_pad * ld);
    if (ns.grid.present) s.dx.ensure((size_t)ns.grid.n_levels * ns.grid.n_feat * ld);
    s.partial.ensure((size_t)n_chunks * ns.n_matrix_params);
}

static int grid_part_index(const NetSpec& ns) {
    for (int i = 0; i < ns.n_parts; ++i)
        if (ns.parts[i].type == ENC_GRID) return i;
    return -1;
}

// grid_done: s.gfeat already holds this batch's grid features (the paired
// theta/phi pre-pass ran).
static void net_forward_train(dfg_ctx* c, NetState& s, const float* feat, int stride, const float* target,
                              int mode, int64_t n, int64_t ld, hipStream_t st, bool grid_done = false) {
    TrainFwdArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ns = s.d_spec.p;
    a.pack = s.pack_train.p;
    a.tab = s.tab_train();
    a.feat = feat;
    a.stride = stride;
    a.target = target;
    a.theta_bins = c->n_theta;
    a.linear = c->linear;
    a.mode = mode;
    a.n = n;
    a.ld = ld;
    a.x = s.x.p;
    a.act = s.act.p;
    a.out = s.out.p;
    a.xs = s.spec.k_steps * 32 + 8;  // halves
    a.ls = s.spec.n_out_pad + 1;
    if (s.spec.grid.present) {
        if (!grid_done) {
            s.gfeat.ensure((size_t)s.spec.grid.n_levels * s.spec.grid.n_feat * ld);
            const float* pos = compact_pos(c->train_pos, feat, stride, grid_pos_dim(s.spec), n, st);
            launch_grid_encode(s.d_spec.p, s.spec, s.tab_train(), pos, n, ld, s.gfeat.p, st);
        }
        a.gfeat = s.gfeat.p;
    }
    launch_train_fwd(a, s.spec, st);
}

// Grid features of both DF nets for a training batch in one XCD-aware pass
// over the interleaved training table (returns false when the grids differ).
static bool train_grid_pair(dfg_ctx* c, const float* feat, int stride, int64_t n, int64_t ld, hipStream_t st) {
    if (!c->pair_ok) return false;
    NetState& th = c->net[0];
    NetState& ph = c->net[1];
    const size_t cols = (size_t)th.spec.grid.n_levels * 2;
    th.gfeat.ensure(cols * ld);
    ph.gfeat.ensure(cols * ld);
    const float* pos = compact_pos(c->train_pos, feat, stride, grid_pos_dim(th.spec), n, st);
    launch_grid_encode_pair(th.d_spec.p, th.spec, c->tab_pair_train.p, pos, n, ld, th.gfeat.p, ph.gfeat.p, st);
    return true;
}

static void net_update(dfg_ctx* c, NetState& s, float loss_scale, hipStream_t st);

// MLP backward of one net: dL/d(activations) and dL/d(grid features) (dx),
// then the weight gradients.
static void net_backward_mlp(NetState& s, int64_t n, int64_t ld, int n_chunks, int chunk, hipStream_t st) {
    const NetSpec& ns = s.spec;
    const int gp = grid_part_index(ns);
    TrainBwdArgs b;
    std::memset(&b, 0, sizeof(b));
    b.ns = s.d_spec.p;
    b.pack = s.pack_train.p;
    b.gout = s.gout_rm.p;
    b.act = s.act.p;
    b.g = s.g.p;
    b.n = n;
    b.ld = ld;
    if (gp >= 0) {
        b.dx = s.dx.p;
        b.dx_begin = ns.parts[gp].out_begin;
        b.dx_n = ns.grid.n_levels * ns.grid.n_feat;
    }
    launch_train_bwd(b, ns, st);
    WgradArgs w;
    std::memset(&w, 0, sizeof(w));
    w.ns = s.d_spec.p;
    w.x = s.x.p;
    w.act = s.act.p;
    w.g = s.g.p;
    w.gout_fm = s.gout_fm.p;
    w.partial = s.partial.p;
    w.grad = s.grad.p;
    w.n = n;
    w.ld = ld;
    w.chunk = chunk;
    w.n_chunks = n_chunks;
    launch_wgrad(w, ns, st);
}

// Arguments of the bucketed grid backward of net s over the context's shared
// bucket buffers, with its records in `rec` and level maxima at `lvl_max`.
static GridBucketArgs bucket_args(dfg_ctx* c, NetState& s, const float* feat, int stride, int64_t n, int64_t ld,
                                  const GridBucketPlan& plan, float* rec, uint32_t* lvl_max) {
    const NetSpec& ns = s.spec;
    GridBucketArgs gb;
    std::memset(&gb, 0, sizeof(gb));
    gb.ns = s.d_spec.p;
    gb.feat = feat;
    gb.stride = stride;
    gb.pos_dim = ns.parts[grid_part_index(ns)].dim_begin;
    gb.dx = s.dx.p;
    gb.grad = s.grad.p;
    gb.n = n;
    gb.ld = ld;
    gb.slice_shift = plan.slice_shift;
    gb.n_chunks = plan.n_chunks;
    gb.offs = c->gb_offs.p;
    gb.cursor = c->gb_cursor.p;
    gb.cnt = c->gb_cnt.p;
    gb.rec_idx = c->gb_idx.p;
    gb.rec_val = rec;
    gb.items = c->gb_items.p;
    gb.n_items = c->gb_n_items.p;
    gb.lvl_max = lvl_max;
    gb.partial = c->gb_part.p;
    gb.multi = c->gb_multi.p;
    return gb;
}

// Shared bucket buffers for a plan with F values per record (offs carries
// `maxima` per-level maxima arrays after the slice offsets).  F == 2: 12-byte
// records in gb_val; paired: 144-byte groups of 8 records in gb_val;
// otherwise an entry array and F values per record.
static void ensure_bucket_ws(dfg_ctx* c, const GridBucketPlan& plan, int F, int maxima, bool paired = false) {
    c->gb_offs.ensure((size_t)plan.n_slices + 1 + (size_t)maxima * DFG_MAX_LEVELS / 2);
    c->gb_part.ensure((size_t)plan.max_partials * kGridSliceFloats);
    c->gb_cursor.ensure((size_t)plan.n_slices);
    c->gb_cnt.ensure((size_t)plan.n_slices * plan.n_chunks);
    if (F != 2 && !paired) c->gb_idx.ensure((size_t)plan.records);
    if (paired)
        c->gb_val.ensure((size_t)(plan.records + 7) / 8 * (kPairGroupBytes / 4));
    else
        c->gb_val.ensure((size_t)plan.records * (F == 2 ? 3 : F));
    c->gb_items.ensure((size_t)plan.max_items);
    c->gb_n_items.ensure(2);
    c->gb_multi.ensure((size_t)plan.max_multi);
}

static void net_grid_backward(dfg_ctx* c, NetState& s, const float* feat, int stride, int64_t n, int64_t ld,
                              hipStream_t st) {
    const NetSpec& ns = s.spec;
    const int gp = grid_part_index(ns);
    if (gp < 0) return;
    const GridBucketPlan plan = grid_bucket_plan(ns, n);
    if (c->grid_bwd_bucketed && plan.ok) {
        ensure_bucket_ws(c, plan, ns.grid.n_feat, 1);
        const GridBucketArgs gb = bucket_args(c, s, feat, stride, n, ld, plan, c->gb_val.p,
                                              (uint32_t*)(c->gb_offs.p + plan.n_slices + 1));
        launch_grid_bwd_bucketed(gb, ns, plan, st);
    } else {
        const size_t ngrid = (size_t)ns.n_params - ns.grid.param_begin;
        launch_fill_f32(s.grad.p + ns.grid.param_begin, 0.0f, (int64_t)ngrid, st);
        GridBwdArgs gb;
        std::memset(&gb, 0, sizeof(gb));
        gb.ns = s.d_spec.p;
        gb.feat = feat;
        gb.stride = stride;
        gb.pos_dim = ns.parts[gp].dim_begin;
        gb.dx = s.dx.p;
        gb.grad = s.grad.p;
        gb.n = n;
        gb.ld = ld;
        launch_grid_bwd(gb, ns, st);
    }
}

// Grid backward of both DF nets in one bucketed pass (interleaved-table
// grids: F = 2, identical structure); false when that does not apply.
// pos: the batch's compact grid positions [n][3] (from the step's paired
// pre-pass), read 16 times (once per level) instead of the strided features.
static bool grid_backward_pair(dfg_ctx* c, const float* feat, int stride, const float* pos, int64_t n, int64_t ld,
                               hipStream_t st, const GridAdam* gadam = nullptr) {
    NetState& th = c->net[0];
    NetState& ph = c->net[1];
    if (!c->pair_ok || !c->grid_bwd_bucketed || th.spec.grid.interp == INTERP_NEAREST) return false;
    // one record per corner for both nets: entry + float4 values, 2,048-entry slices
    const GridBucketPlan plan = grid_bucket_plan(th.spec, n, 4);
    if (!plan.ok) return false;
    ensure_bucket_ws(c, plan, 4, 2, true);  // groups of 8 records: 8 float4 values + 8 16-bit entries
    uint32_t* mx = (uint32_t*)(c->gb_offs.p + plan.n_slices + 1);
    GridBucketArgs a = bucket_args(c, th, feat, stride, n, ld, plan, c->gb_val.p, mx);
    if (pos) {
        a.feat = pos;
        a.stride = 3;
        a.pos_dim = 0;
    }
    a.dx2 = ph.dx.p;
    a.grad2 = ph.grad.p;
    a.param_begin2 = ph.spec.grid.param_begin;
    a.lvl_max2 = mx + DFG_MAX_LEVELS;
    GridAdam off;
    std::memset(&off, 0, sizeof(off));
    launch_grid_bwd_bucketed_pair(a, th.spec, plan, gadam ? *gadam : off, st);
    return true;
}

static void net_backward_update(dfg_ctx* c, NetState& s, const float* feat, int stride, int64_t n, int64_t ld,
                                int n_chunks, int chunk, float loss_scale, hipStream_t st) {
    net_backward_mlp(s, n, ld, n_chunks, chunk, st);
    net_grid_backward(c, s, feat, stride, n, ld, st);
    net_update(c, s, loss_scale, st);
}

// Adam + EMA (EmaOptimizer::step, ema.h:102-136) on the gradient in s.grad,
// then the fragment packs of the new weights.
// Adam + EMA arguments of net s for its next step (advances its step count).
static AdamArgs adam_args(dfg_ctx* c, NetState& s, float loss_scale) {
    const NetSpec& ns = s.spec;
    s.step += 1;
    AdamArgs o;
    std::memset(&o, 0, sizeof(o));
    o.n_params = ns.n_params;
    o.n_matrix = ns.n_matrix_params;
    o.lr = s.opt.lr;
    o.beta1 = s.opt.beta1;
    o.beta2 = s.opt.beta2;
    o.log2_beta1 = (float)std::log2((double)s.opt.beta1);
    o.log2_beta2 = (float)std::log2((double)s.opt.beta2);
    const int slot = (int)(&s - c->net);
    if (c->pair_ok && slot < 2) {
        o.pair_w16 = (_Float16*)c->tab_pair_train.p;
        o.pair_ema16 = (_Float16*)c->tab_pair.p;
        o.grid_begin = ns.grid.param_begin;
        o.pair_slot = slot;
    }
    o.eps = s.opt.eps;
    o.l2_reg = s.opt.l2;
    o.loss_scale = loss_scale;
    o.ema_decay = s.opt.ema_decay;
    o.ema_debias_old = 1.0f - (float)std::pow((double)s.opt.ema_decay, (double)(s.step - 1));
    o.ema_debias_new = 1.0f / (1.0f - (float)std::pow((double)s.opt.ema_decay, (double)s.step));
    o.master = s.master.p;
    o.w16 = s.w16.p;
    o.ema16 = s.ema16.p;
    o.grad = s.grad.p;
    o.m = s.m.p;
    o.v = s.v.p;
    o.steps = s.short_steps ? nullptr : s.steps.p;
    o.steps16 = s.short_steps ? s.steps16.p : nullptr;
    return o;
}

// The training (forward + backward) and inference (forward, EMA) packs of the
// given nets, in one launch.
static void nets_pack(std::initializer_list<NetState*> nets, hipStream_t st) {
    PackJobs jobs;
    std::memset(&jobs, 0, sizeof(jobs));
    for (NetState* s : nets) {
        jobs.job[jobs.n++] = PackJob{s->d_spec.p, s->w16.p, s->pack_train.p, 3, pack_slots(s->spec, 3)};
        jobs.job[jobs.n++] = PackJob{s->d_spec.p, s->ema16.p, s->pack_infer.p, 1, pack_slots(s->spec, 1)};
    }
    launch_pack_jobs(jobs, st);
}

static void net_pack(NetState& s, hipStream_t st) { nets_pack({&s}, st); }

static void net_update(dfg_ctx* c, NetState& s, float loss_scale, hipStream_t st) {
    s.grad_kept = true;
    launch_adam_ema(adam_args(c, s, loss_scale), st);
    net_pack(s, st);
}

// Adam + EMA of both DF nets: each net's MLP parameters on their own, the
// grid parameters of the two nets together (whole interleaved-table entries).
static void nets_update_pair(dfg_ctx* c, float loss_scale, hipStream_t st) {
    NetState& th = c->net[0];
    NetState& ph = c->net[1];
    const int64_t entries = (int64_t)th.spec.grid.offset[th.spec.grid.n_levels];
    // the grid parameters: the whole tail of each net's parameters, past the
    // matrix parameters, starting at an even index (8-byte vector accesses)
    auto tail = [&](const NetState& s) {
        const int64_t b = s.spec.grid.param_begin;
        return b + 2 * entries == s.spec.n_params && b >= s.spec.n_matrix_params && b % 2 == 0;
    };
    const bool whole = c->pair_ok && tail(th) && tail(ph);
    if (!whole) {
        net_update(c, th, loss_scale, st);
        net_update(c, ph, loss_scale, st);
        return;
    }
    const AdamArgs a = adam_args(c, th, loss_scale);
    const AdamArgs b = adam_args(c, ph, loss_scale);
    th.grad_kept = ph.grad_kept = true;
    launch_adam_ema_pair(a, b, entries, st);
    nets_pack({&th, &ph}, st);
}

// Grid backward of both DF nets with their grids' Adam + EMA inside it (each
// slice's sums update its entries; no gradient round trip through HBM), then
// the MLP parameters' Adam + EMA and the packs.  False (nothing launched, no
// step counted) when the paired bucketed backward does not apply.
static bool grid_backward_update_pair(dfg_ctx* c, const float* feat, int stride, const float* pos, int64_t n,
                                      int64_t ld, float loss_scale, hipStream_t st) {
    NetState& th = c->net[0];
    NetState& ph = c->net[1];
    if (!c->pair_ok || !c->grid_bwd_bucketed || th.spec.grid.interp == INTERP_NEAREST) return false;
    const int64_t entries = (int64_t)th.spec.grid.offset[th.spec.grid.n_levels];
    auto tail = [&](const NetState& s) {
        const int64_t b = s.spec.grid.param_begin;
        return b + 2 * entries == s.spec.n_params && b >= s.spec.n_matrix_params && b % 2 == 0;
    };
    if (!tail(th) || !tail(ph) || !grid_bucket_plan(th.spec, n, 4).ok) return false;
    GridAdam ga;
    std::memset(&ga, 0, sizeof(ga));
    ga.a = adam_args(c, th, loss_scale);
    ga.b = adam_args(c, ph, loss_scale);
    ga.on = 1;
    ga.keep_grad = c->keep_gradients ? 1 : 0;
    if (!grid_backward_pair(c, feat, stride, pos, n, ld, st, &ga))
        throw Error(DFG_ERR_HIP, "paired grid backward refused a plan it accepted");
    launch_adam_ema_pair(ga.a, ga.b, 0, st);  // the MLP parameters of both nets
    nets_pack({&th, &ph}, st);
    th.grad_kept = ph.grad_kept = c->keep_gradients;
    return true;
}

static const int kChunk = 1024;

int dfg_train_step(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride, const float* target,
                   float* loss_out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, B, feat, feat_stride);
    if (B > 0 && !target) throw Error(DFG_ERR_INVALID_ARG, "null target");
    if (B == 0) return DFG_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t ld = round_up(B, 64);
    const int chunk = ctx->wgrad_chunk_rows;
    const int n_chunks = (int)((ld + chunk - 1) / chunk);
    NetState& th = ctx->net[0];
    NetState& ph = ctx->net[1];
    ctx->loss.ensure((size_t)ld);
    if (ctx->fused_training && train_fused_df_ok(th.spec, ctx->n_theta) && train_fused_df_ok(ph.spec, ctx->n_phi)) {
        // per net one pass over 64-row tiles kept in LDS (forward, loss,
        // backward, weight gradients); the loss couples the nets through each
        // row's |pdf|: phi's forward alone first, then theta's and phi's passes
        const int64_t n_tiles = ld / 64;
        const int blocks = (int)std::min<int64_t>(n_tiles, 2 * ctx->n_cus);
        th.partial.ensure((size_t)blocks * th.spec.n_matrix_params);
        ph.partial.ensure((size_t)blocks * ph.spec.n_matrix_params);
        ctx->train_pdf.ensure((size_t)2 * ld);
        float* pdf_th = ctx->train_pdf.p;
        float* pdf_ph = ctx->train_pdf.p + ld;
        // the grid features of the batch: one pass over the interleaved table
        // for grids of identical structure, else one pre-pass per net
        for (NetState* s : {&th, &ph})
            if (s->spec.grid.present) s->dx.ensure((size_t)s->spec.grid.n_levels * s->spec.grid.n_feat * ld);
        if (!train_grid_pair(ctx, feat, feat_stride, B, ld, st)) {
            for (NetState* s : {&th, &ph}) {
                if (!s->spec.grid.present) continue;
                s->gfeat.ensure((size_t)s->spec.grid.n_levels * s->spec.grid.n_feat * ld);
                const float* pos = compact_pos(ctx->train_pos, feat, feat_stride, grid_pos_dim(s->spec), B, st);
                launch_grid_encode(s->d_spec.p, s->spec, s->tab_train(), pos, B, ld, s->gfeat.p, st);
            }
        }
        auto args = [&](NetState& s, int theta_in) {
            TrainFusedArgs f;
            std::memset(&f, 0, sizeof(f));
            f.ns = s.d_spec.p;
            f.pack = s.pack_train.p;
            f.feat = feat;
            f.stride = feat_stride;
            f.target = target;
            f.loss = ctx->loss.p;
            f.partial = s.partial.p;
            f.n = B;
            f.n_tiles = n_tiles;
            f.loss_scale = 1024.0f;  // TRAIN_LOSS_SCALE, parameters.h:45
            f.ld = ld;
            f.theta_in = theta_in;
            f.theta_bins = ctx->n_theta;
            f.linear = ctx->linear;
            const int gp = grid_part_index(s.spec);
            if (gp >= 0) {
                f.gfeat = s.gfeat.p;
                f.dx = s.dx.p;
                f.dx_begin = s.spec.parts[gp].out_begin;
                f.dx_n = s.spec.grid.n_levels * s.spec.grid.n_feat;
            }
            return f;
        };
        TrainFusedArgs fp = args(ph, 1);
        fp.pdf_out = pdf_ph;
        TrainFusedArgs ft = args(th, 0);
        ft.pdf_in = pdf_ph;
        ft.pdf_out = pdf_th;
        TrainFusedArgs fq = args(ph, 1);
        fq.pdf_in = pdf_th;
        // the forward-only pass holds no activation slots: three workgroups per CU
        const int blocks_pdf = (int)std::min<int64_t>(n_tiles, 3 * ctx->n_cus);
        if (!launch_train_fused_df(fp, ph.spec, ctx->n_phi, true, blocks_pdf, st) ||
            !launch_train_fused_df(ft, th.spec, ctx->n_theta, false, blocks, st) ||
            !launch_train_fused_df(fq, ph.spec, ctx->n_phi, false, blocks, st))
            throw Error(DFG_ERR_HIP, "fused DF training launch failed");
        launch_wgrad_reduce(th.partial.p, blocks, (int64_t)th.spec.n_matrix_params, th.grad.p, st);
        launch_wgrad_reduce(ph.partial.p, blocks, (int64_t)ph.spec.n_matrix_params, ph.grad.p, st);
        if (!grid_backward_update_pair(ctx, feat, feat_stride, ctx->train_pos.p, B, ld, 1024.0f, st)) {
            if (!grid_backward_pair(ctx, feat, feat_stride, ctx->train_pos.p, B, ld, st)) {
                net_grid_backward(ctx, th, feat, feat_stride, B, ld, st);
                net_grid_backward(ctx, ph, feat, feat_stride, B, ld, st);
            }
            nets_update_pair(ctx, 1024.0f, st);
        }
        if (loss_out) {
            ctx->mean_ws.ensure(256);
            launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
        }
        check_launch();
        return DFG_OK;
    }
    ensure_train_ws(th, ld, n_chunks);
    ensure_train_ws(ph, ld, n_chunks);
    const bool paired = train_grid_pair(ctx, feat, feat_stride, B, ld, st);
    net_forward_train(ctx, th, feat, feat_stride, nullptr, 0, B, ld, st, paired);
    net_forward_train(ctx, ph, feat, feat_stride, target, 0, B, ld, st, paired);
    DfLossArgs l;
    std::memset(&l, 0, sizeof(l));
    l.target = target;
    l.bins_th = th.out.p;
    l.bins_ph = ph.out.p;
    l.nt_pad = th.spec.n_out_pad;
    l.np_pad = ph.spec.n_out_pad;
    l.gout_th_rm = th.gout_rm.p;
    l.gout_ph_rm = ph.gout_rm.p;
    l.gout_th_fm = th.gout_fm.p;
    l.gout_ph_fm = ph.gout_fm.p;
    l.loss = ctx->loss.p;
    l.n = B;
    l.ld = ld;
    l.loss_scale = 1024.0f;  // TRAIN_LOSS_SCALE, parameters.h:45
    l.linear = ctx->linear;
    launch_df_loss(l, ctx->n_theta, ctx->n_phi, st);
    if (paired) {
        // both MLP backwards, then one grid backward for the two grids
        net_backward_mlp(th, B, ld, n_chunks, chunk, st);
        net_backward_mlp(ph, B, ld, n_chunks, chunk, st);
        // the compact positions of train_grid_pair are still in ctx->train_pos
        if (!grid_backward_update_pair(ctx, feat, feat_stride, ctx->train_pos.p, B, ld, 1024.0f, st)) {
            if (!grid_backward_pair(ctx, feat, feat_stride, ctx->train_pos.p, B, ld, st)) {
                net_grid_backward(ctx, th, feat, feat_stride, B, ld, st);
                net_grid_backward(ctx, ph, feat, feat_stride, B, ld, st);
            }
            nets_update_pair(ctx, 1024.0f, st);
        }
    } else {
        net_backward_update(ctx, th, feat, feat_stride, B, ld, n_chunks, chunk, 1024.0f, st);
        net_backward_update(ctx, ph, feat, feat_stride, B, ld, n_chunks, chunk, 1024.0f, st);
    }
    if (loss_out) {
        ctx->mean_ws.ensure(256);
        launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
    }
    check_launch();
    API_END(ctx)
}

int dfg_nrc_train_step(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                       const float* target_rgb, float* loss_out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (B < 0 || B > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "B out of range");
    if (feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below 15");
    if (B > 0 && (!feat || !target_rgb)) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (B == 0) return DFG_OK;
    const NrcScratch scope(ctx);  // its own step scratch: may run beside a DF step on another stream
    hipStream_t st = (hipStream_t)stream;
    const int64_t ld = round_up(B, 64);
    const int chunk = ctx->wgrad_chunk_rows;
    const int n_chunks = (int)((ld + chunk - 1) / chunk);
    NetState& s = ctx->net[2];
    ctx->loss.ensure((size_t)ld);
    if (ctx->fused_training) {
        // forward, loss, backward and weight gradients in one pass over
        // 64-row tiles kept in LDS, two workgroups per CU
        const int64_t n_tiles = ld / 64;
        const int blocks = (int)std::min<int64_t>(n_tiles, 2 * ctx->n_cus);
        s.partial.ensure((size_t)blocks * s.spec.n_matrix_params);
        TrainFusedArgs f;
        std::memset(&f, 0, sizeof(f));
        f.ns = s.d_spec.p;
        f.pack = s.pack_train.p;
        f.tab = s.tab_train();
        f.feat = feat;
        f.stride = feat_stride;
        f.target = target_rgb;
        f.loss = ctx->loss.p;
        f.partial = s.partial.p;
        f.n = B;
        f.n_tiles = n_tiles;
        f.loss_scale = 1024.0f;  // TRAIN_LOSS_SCALE_NRC, parameters.h:59
        if (launch_train_fused_nrc(f, s.spec, blocks, st)) {
            launch_wgrad_reduce(s.partial.p, blocks, (int64_t)s.spec.n_matrix_params, s.grad.p, st);
            net_update(ctx, s, 1024.0f, st);
            if (loss_out) {
                ctx->mean_ws.ensure(256);
                launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
            }
            check_launch();
            return DFG_OK;
        }
    }
    ensure_train_ws(s, ld, n_chunks);
    net_forward_train(ctx, s, feat, feat_stride, nullptr, 1, B, ld, st);
    NrcLossArgs l;
    std::memset(&l, 0, sizeof(l));
    l.pred = s.out.p;
    l.target = target_rgb;
    l.gout_rm = s.gout_rm.p;
    l.gout_fm = s.gout_fm.p;
    l.loss = ctx->loss.p;
    l.n = B;
    l.ld = ld;
    l.loss_scale = 1024.0f;  // TRAIN_LOSS_SCALE_NRC, parameters.h:59
    launch_nrc_loss(l, st);
    net_backward_update(ctx, s, feat, feat_stride, B, ld, n_chunks, chunk, 1024.0f, st);
    if (loss_out) {
        ctx->mean_ws.ensure(256);
        launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
    }
    check_launch();
    API_END(ctx)
}

// --- plain FP32 training step (comparator) ---------------------------------------------
static void ensure_plain_train_ws(NetState& s, int64_t ld, int n_chunks) {
    const NetSpec& ns = s.spec;
    s.px.ensure((size_t)ns.enc_width * ld);
    s.pact.ensure((size_t)std::max(ns.n_hidden, 1) * 64 * ld);
    s.pout.ensure((size_t)ns.n_out_pad * ld);
    s.pgout.ensure((size_t)ns.n_out_pad * ld);
    s.pg.ensure((size_t)2 * 64 * ld);
    if (ns.grid.present) s.dx.ensure((size_t)ns.grid.n_levels * ns.grid.n_feat * ld);
    s.partial.ensure((size_t)n_chunks * ns.n_matrix_params);
}

// encode, then one fp32 kernel per layer; hidden activations kept for the backward
static void plain_net_forward(NetState& s, const float* feat, int stride, const float* extra, int64_t n, int64_t ld,
                              hipStream_t st) {
    const NetSpec& ns = s.spec;
    launch_plain_encode(s.d_spec.p, s.tab_train(), feat, stride, extra, n, ld, s.px.p, st);
    const float* X = s.px.p;
    for (int l = 0; l <= ns.n_hidden; ++l) {
        const int M = l == ns.n_hidden ? ns.n_out_pad : 64;
        const int K = l == 0 ? ns.enc_width : 64;
        float* Y = l == ns.n_hidden ? s.pout.p : s.pact.p + (size_t)l * 64 * ld;
        launch_plain_layer(s.w16.p + ns.layer_off[l], M, K, K, X, Y, l < ns.n_hidden, n, ld, st);
        X = Y;
    }
}

// from the loss-scaled output gradient in s.pgout: per layer the weight
// gradient and the ReLU-masked data gradient, the grid features' gradient,
// the grid scatter, then Adam + EMA
static void plain_net_backward_update(dfg_ctx* c, NetState& s, const float* feat, int stride, int64_t n, int64_t ld,
                                      int n_chunks, int chunk, float loss_scale, hipStream_t st) {
    const NetSpec& ns = s.spec;
    const int H = ns.n_hidden;
    const int gp = grid_part_index(ns);
    const float* dY = s.pgout.p;
    for (int l = H; l >= 0; --l) {
        const int M = l == H ? ns.n_out_pad : 64;  // outputs of layer l
        const int K = l == 0 ? ns.enc_width : 64;  // inputs of layer l
        const float* in = l == 0 ? s.px.p : s.pact.p + (size_t)(l - 1) * 64 * ld;
        launch_plain_wgrad(dY, M, in, K, s.partial.p + ns.layer_off[l], ns.n_matrix_params, n, ld, chunk, n_chunks,
                           st);
        if (l > 0) {
            float* dX = s.pg.p + (size_t)(l & 1) * 64 * ld;
            launch_plain_layer_bwd(s.w16.p + ns.layer_off[l], K, 0, 64, M, dY, in, dX, n, ld, st);
            dY = dX;
        } else if (gp >= 0) {
            const int nf = ns.grid.n_levels * ns.grid.n_feat;
            if (nf > 64) throw Error(DFG_ERR_CONFIG, "plain path: more than 64 grid features");
            launch_plain_layer_bwd(s.w16.p + ns.layer_off[0], K, ns.parts[gp].out_begin, nf, M, dY, nullptr, s.dx.p,
                                   n, ld, st);
        }
    }
    launch_wgrad_reduce(s.partial.p, n_chunks, ns.n_matrix_params, s.grad.p, st);
    if (gp >= 0) {
        launch_fill_f32(s.grad.p + ns.grid.param_begin, 0.0f, (int64_t)ns.n_params - ns.grid.param_begin, st);
        GridBwdArgs gb;
        std::memset(&gb, 0, sizeof(gb));
        gb.ns = s.d_spec.p;
        gb.feat = feat;
        gb.stride = stride;
        gb.pos_dim = ns.parts[gp].dim_begin;
        gb.dx = s.dx.p;
        gb.grad = s.grad.p;
        gb.n = n;
        gb.ld = ld;
        launch_grid_bwd(gb, ns, st);
    }
    net_update(c, s, loss_scale, st);
}

int dfg_train_step_plain(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                         const float* target, float* loss_out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    check_rows(ctx, B, feat, feat_stride);
    if (B > 0 && !target) throw Error(DFG_ERR_INVALID_ARG, "null target");
    if (B == 0) return DFG_OK;
    hipStream_t st = (hipStream_t)stream;
    // the reference synchronises the device at the start of every training batch
    HIPX(hipStreamSynchronize(st));
    const int64_t ld = round_up(B, 64);
    const int n_chunks = (int)((ld + kChunk - 1) / kChunk);
    NetState& th = ctx->net[0];
    NetState& ph = ctx->net[1];
    ensure_plain_train_ws(th, ld, n_chunks);
    ensure_plain_train_ws(ph, ld, n_chunks);
    ctx->loss.ensure((size_t)ld);
    ctx->p_theta.ensure((size_t)ld);
    plain_net_forward(th, feat, feat_stride, nullptr, B, ld, st);
    launch_plain_theta_target(target, B, ctx->n_theta, ctx->linear, ctx->p_theta.p, st);
    plain_net_forward(ph, feat, feat_stride, ctx->p_theta.p, B, ld, st);
    launch_plain_df_loss(target, th.pout.p, ph.pout.p, ctx->n_theta, ctx->n_phi, ctx->linear, th.pgout.p, ph.pgout.p,
                         ctx->loss.p, 1024.0f, B, ld, st);
    plain_net_backward_update(ctx, th, feat, feat_stride, B, ld, n_chunks, kChunk, 1024.0f, st);
    plain_net_backward_update(ctx, ph, feat, feat_stride, B, ld, n_chunks, kChunk, 1024.0f, st);
    if (loss_out) {
        ctx->mean_ws.ensure(256);
        launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
    }
    check_launch();
    API_END(ctx)
}

int dfg_nrc_train_step_plain(dfg_ctx* ctx, void* stream, int64_t B, const float* feat, int feat_stride,
                             const float* target_rgb, float* loss_out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    if (!ctx->net[2].present) throw Error(DFG_ERR_INVALID_ARG, "no NRC net configured");
    if (B < 0 || B > ((int64_t)1 << 30)) throw Error(DFG_ERR_BAD_SIZE, "B out of range");
    if (feat_stride < 15) throw Error(DFG_ERR_BAD_SIZE, "feat_stride below 15");
    if (B > 0 && (!feat || !target_rgb)) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    if (B == 0) return DFG_OK;
    const NrcScratch scope(ctx);  // its own step scratch: may run beside a DF step on another stream
    hipStream_t st = (hipStream_t)stream;
    HIPX(hipStreamSynchronize(st));
    const int64_t ld = round_up(B, 64);
    const int n_chunks = (int)((ld + kChunk - 1) / kChunk);
    NetState& s = ctx->net[2];
    ensure_plain_train_ws(s, ld, n_chunks);
    ctx->loss.ensure((size_t)ld);
    plain_net_forward(s, feat, feat_stride, nullptr, B, ld, st);
    launch_plain_nrc_loss(s.pout.p, target_rgb, s.pgout.p, ctx->loss.p, 1024.0f, B, ld, st);
    plain_net_backward_update(ctx, s, feat, feat_stride, B, ld, n_chunks, kChunk, 1024.0f, st);
    if (loss_out) {
        ctx->mean_ws.ensure(256);
        launch_mean(ctx->loss.p, B, 1.0f / (float)B, loss_out, ctx->mean_ws.p, st);
    }
    check_launch();
    API_END(ctx)
}

// --- stateless & debug ---------------------------------------------------------------
int dfg_prepare_features(void* stream, int64_t n, const float* pos, const float* aabb, const float* wo,
                         const float* normal, const float* roughness, float* feat) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    if (n == 0) return DFG_OK;
    if (!pos || !aabb || !wo || !normal || !roughness || !feat) return DFG_ERR_INVALID_ARG;
    launch_prepare_features(n, pos, aabb, wo, normal, roughness, feat, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

int dfg_df_head(void* stream, int64_t n, int n_theta, int n_phi, int interpolation, const float* logits_theta,
                const float* logits_phi, const float* u, const uint8_t* guided, float* wi, float* bins_theta,
                float* bins_phi, int32_t* seg) {
    if (n < 0) return DFG_ERR_BAD_SIZE;
    auto ok = [](int b) { return b == 16 || b == 32 || b == 64; };
    if (!ok(n_theta) || !ok(n_phi)) return DFG_ERR_INVALID_ARG;
    if (n == 0) return DFG_OK;
    if (!logits_theta || !logits_phi || !u || !wi) return DFG_ERR_INVALID_ARG;
    launch_df_head(n, n_theta, n_phi, interpolation ? 1 : 0, logits_theta, logits_phi, u, guided, wi, bins_theta,
                   bins_phi, seg, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? DFG_OK : DFG_ERR_HIP;
}

int dfg_encode(dfg_ctx* ctx, void* stream, int net, int64_t n, const float* in, int in_stride, float* out) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    NetState& s = need_net(ctx, net);
    if (n < 0) throw Error(DFG_ERR_BAD_SIZE, "n < 0");
    if (in_stride < s.spec.in_dims) throw Error(DFG_ERR_BAD_SIZE, "in_stride below the net's input dims");
    if (n == 0) return DFG_OK;
    if (!in || !out) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    launch_encode_debug(s.d_spec.p, s.spec.enc_width, s.tab_infer(), in, in_stride, n, out, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

int dfg_grid_index(dfg_ctx* ctx, void* stream, int net, int64_t n, const float* in, int in_stride, uint32_t* idx) {
    if (!ctx) return DFG_ERR_NOT_INIT;
    API_BEGIN
    NetState& s = need_net(ctx, net);
    if (!s.spec.grid.present) throw Error(DFG_ERR_INVALID_ARG, "net has no grid encoding");
    if (n < 0) throw Error(DFG_ERR_BAD_SIZE, "n < 0");
    if (in_stride < s.spec.in_dims) throw Error(DFG_ERR_BAD_SIZE, "in_stride below the net's input dims");
    if (n == 0) return DFG_OK;
    if (!in || !idx) throw Error(DFG_ERR_INVALID_ARG, "null argument");
    launch_grid_index_debug(s.d_spec.p, in, in_stride, n, idx, (hipStream_t)stream);
    check_launch();
    API_END(ctx)
}

}  // extern "C"
