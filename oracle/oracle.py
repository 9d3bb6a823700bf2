"""ORACLE — test infrastructure only (tests/, the smoke check and the bench's
CPU-baseline leg use it as the checker; the product never imports it).

Python half of the CPU restatement of the reference's DF guiding math:
  * layout(): an independent derivation of tiny-cuda-nn's encoding layout
    (composite padding, encodings/composite.h:136-215, encoding.h:85-87;
    grid level offsets, encodings/grid.h:959-1025) and of the FullyFusedMLP
    parameter layout (src/fully_fused_mlp.cu:636-688,
    network_with_input_encoding.h:113-155);
  * numpy wrappers around liboracle.so (dfg_oracle.cpp: FP64 oracle and the
    FP32 reference-order mode).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ENC = {"identity": 0, "oneblob": 1, "sphericalharmonics": 2, "grid": 3, "hashgrid": 3, "densegrid": 3,
       "tiledgrid": 3, "trianglewave": 4, "frequency": 5}
INTERP = {"nearest": 0, "linear": 1, "smoothstep": 2}
GRID_TYPE = {"dense": 0, "hash": 1, "tiled": 2}


class OrcPart(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("type", "dim_begin", "n_dims", "out_begin", "n_out", "n_pad", "pad_front", "param")] + \
               [("id_scale", ctypes.c_float), ("id_offset", ctypes.c_float)]


class OrcNet(ctypes.Structure):
    _fields_ = [("in_dims", ctypes.c_int), ("n_parts", ctypes.c_int), ("parts", OrcPart * 6),
                ("enc_width", ctypes.c_int), ("n_hidden", ctypes.c_int), ("n_out", ctypes.c_int),
                ("n_out_pad", ctypes.c_int), ("grid_present", ctypes.c_int), ("grid_levels", ctypes.c_int),
                ("grid_F", ctypes.c_int), ("grid_interp", ctypes.c_int), ("grid_type", ctypes.c_int),
                ("grid_offset", ctypes.c_uint32 * 25), ("grid_res", ctypes.c_uint32 * 24),
                ("grid_scale", ctypes.c_float * 24), ("layer_off", ctypes.c_uint32 * 10),
                ("n_matrix", ctypes.c_uint32), ("n_params", ctypes.c_uint32),
                ("grid_param_begin", ctypes.c_uint32)]


def _next_multiple(v, m):
    return (v + m - 1) // m * m


_libm = ctypes.CDLL("libm.so.6")
_libm.exp2f.argtypes = [ctypes.c_float]
_libm.exp2f.restype = ctypes.c_float
_libm.log2f.argtypes = [ctypes.c_float]
_libm.log2f.restype = ctypes.c_float


def grid_levels(n_levels, base_resolution, per_level_scale, log2_hashmap_size, grid_type):
    """Per-level (scale, resolution, params_in_level) of tcnn's GridEncodingTemplated
    constructor for 3-D positions (grid.h:984-1013), in fp32 like the reference:
    scale = exp2f(level * log2f(s)) * base - 1 (grid.h:194-203)."""
    l2s = np.float32(_libm.log2f(float(np.float32(per_level_scale))))
    out = []
    for l in range(n_levels):
        e = np.float32(_libm.exp2f(float(np.float32(l) * l2s)))
        scale = np.float32(e * np.float32(base_resolution)) - np.float32(1.0)
        scale = np.float32(scale)
        res = int(math.ceil(float(scale))) + 1
        max_params = 0xffffffff // 2
        params = max_params if float(res) ** 3 > max_params else res ** 3
        params = _next_multiple(params, 8)
        if grid_type == "tiled":
            params = min(params, base_resolution ** 3)
        elif grid_type == "hash":
            params = min(params, 1 << log2_hashmap_size)
        out.append((scale, res, params))
    return out


def layout(enc, net, in_dims, n_out):
    """Layout of one NetworkWithInputEncoding(in_dims -> n_out) as a dict."""
    nested = enc["nested"] if enc.get("otype", "OneBlob").lower() == "composite" else [enc]
    specified = sum(e.get("n_dims_to_encode", 0) for e in nested)
    parts = []
    off = 0
    for e in nested:
        dims = e["n_dims_to_encode"] if "n_dims_to_encode" in e else in_dims - specified
        off = e.get("dims_to_encode_begin", off)
        if dims > 0:
            parts.append({"cfg": e, "type": ENC[e.get("otype", "OneBlob").lower()], "n_dims": dims,
                          "dim_begin": off})
        off += dims
    for p in parts:
        e, t, d = p["cfg"], p["type"], p["n_dims"]
        if t == 0:
            p["param"], p["width"] = 0, d
        elif t == 1:
            p["param"] = e.get("n_bins", 16)
            p["width"] = d * p["param"]
        elif t == 2:
            p["param"] = e.get("degree", 4)
            p["width"] = p["param"] ** 2
        elif t == 4:
            p["param"] = e.get("n_frequencies", 12)
            p["width"] = d * p["param"]
        elif t == 5:
            p["param"] = e.get("n_frequencies", 12)
            p["width"] = 2 * d * p["param"]
        elif t == 3:
            p["param"] = e.get("n_features_per_level", 2)
            p["width"] = e.get("n_levels", 16) * p["param"]
        p["align"] = p["param"] if t == 3 else 1
    total_align = 16
    for p in parts:
        total_align = total_align * p["align"] // math.gcd(total_align, p["align"])
    so_far = 0
    for i, p in enumerate(parts):
        a = parts[i + 1]["align"] if i + 1 < len(parts) else total_align
        padded = _next_multiple(so_far + p["width"], a) - so_far
        p["out_begin"], p["n_pad"] = so_far, padded - p["width"]
        so_far += padded
    L = {"in_dims": in_dims, "parts": parts, "enc_width": so_far,
         "n_hidden": net.get("n_hidden_layers", 3), "n_out": n_out, "n_out_pad": _next_multiple(n_out, 16)}
    offs, o = [], 0
    for l in range(L["n_hidden"] + 1):
        rows = L["n_out_pad"] if l == L["n_hidden"] else 64
        cols = L["enc_width"] if l == 0 else 64
        offs.append(o)
        o += rows * cols
    offs.append(o)
    L["layer_off"], L["n_matrix"] = offs, o
    L["grid"] = None
    for p in parts:
        if p["type"] != 3:
            continue
        e = p["cfg"]
        ot = e.get("otype", "Grid").lower()
        gtype = e.get("type", "Tiled" if ot == "tiledgrid" else ("Dense" if ot == "densegrid" else "Hash")).lower()
        lv = grid_levels(e.get("n_levels", 16), e.get("base_resolution", 16), e.get("per_level_scale", 2.0),
                         e.get("log2_hashmap_size", 19), gtype)
        offsets = [0]
        for (_, _, pl) in lv:
            offsets.append(offsets[-1] + pl)
        L["grid"] = {"type": gtype, "F": p["param"], "levels": len(lv), "scale": [s for s, _, _ in lv],
                     "res": [r for _, r, _ in lv], "offset": offsets,
                     "interp": e.get("interpolation", "Linear").lower(), "param_begin": o, "dim_begin": p["dim_begin"],
                     "out_begin": p["out_begin"]}
    L["n_params"] = o + (L["grid"]["offset"][-1] * L["grid"]["F"] if L["grid"] else 0)
    return L


def config_layouts(cfg):
    """(theta, phi, nrc-or-None) layouts of a full guide configuration dict."""
    d = cfg.get("dfg", {})
    nt, npb, D = d.get("n_bins_theta", 16), d.get("n_bins_phi", 32), d.get("n_input_dims", 9)
    nn = cfg["nn"]
    pick = lambda a, b: nn[a] if a in nn else nn[b]
    th = layout(pick("theta_encoding", "encoding"), pick("theta_network", "network"), D, nt)
    ph = layout(pick("phi_encoding", "encoding"), pick("phi_network", "network"), D + 1, npb)
    nrc = None
    if "nrc" in cfg:
        nrc = layout(cfg["nrc"]["encoding"], cfg["nrc"]["network"], 15, 3)
    return th, ph, nrc


def to_struct(L):
    s = OrcNet()
    s.in_dims = L["in_dims"]
    s.n_parts = len(L["parts"])
    for i, p in enumerate(L["parts"]):
        q = s.parts[i]
        q.type, q.dim_begin, q.n_dims = p["type"], p["dim_begin"], p["n_dims"]
        q.out_begin, q.n_out, q.n_pad = p["out_begin"], p["width"], p["n_pad"]
        q.pad_front = 1 if p["type"] == 2 else 0
        q.param = p["param"]
        q.id_scale = p["cfg"].get("scale", 1.0)
        q.id_offset = p["cfg"].get("offset", 0.0)
    s.enc_width, s.n_hidden, s.n_out, s.n_out_pad = L["enc_width"], L["n_hidden"], L["n_out"], L["n_out_pad"]
    for i, o in enumerate(L["layer_off"]):
        s.layer_off[i] = o
    s.n_matrix, s.n_params = L["n_matrix"], L["n_params"]
    g = L["grid"]
    if g:
        s.grid_present, s.grid_levels, s.grid_F = 1, g["levels"], g["F"]
        s.grid_interp = INTERP[g["interp"]]
        s.grid_type = GRID_TYPE[g["type"]]
        for i, o in enumerate(g["offset"]):
            s.grid_offset[i] = o
        for i in range(g["levels"]):
            s.grid_res[i] = g["res"][i]
            s.grid_scale[i] = float(g["scale"][i])
        s.grid_param_begin = g["param_begin"]
    else:
        s.grid_param_begin = L["n_params"]
    return s


# ----------------------------------------------------------------------------
# liboracle.so
# ----------------------------------------------------------------------------
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        F = ctypes.c_float
        U32 = ctypes.c_uint32
        sig = {
            "orc_det_expf": ([F], F), "orc_det_logf": ([F], F), "orc_half_round": ([F], F),
            "orc_grid_index": ([P, P, I64, I, I, P], None),
            "orc_grid_index3": ([I, U32, U32, U32, U32, U32], U32),
            "orc_encode_f64": ([P, P, P, I64, I, P], None),
            "orc_mlp_f64": ([P, P, P, I64, P], None),
            "orc_head_f64": ([P, I64, I, I, P, P, P, P, P, P, P, P], None),
            "orc_head_f32": ([P, I64, I, I, P, P, P, P, P, P, P, P], None),
            "orc_pdf_linear_f64": ([P, I, D], D),
            "orc_theta_input_f64": ([D, I, I], D),
            "orc_theta_input_f32": ([F, I, I], F),
            "orc_query_f64": ([P, P, P, P, P, I, I, I64, P, P, P, I, P, P], None),
            "orc_query_f32_omp": ([P, P, P, P, P, I, I, I64, P, P, P, I, P], I),
            "orc_nrc_f64": ([P, P, P, I, I64, P], None),
            "orc_nrc_f32_omp": ([P, P, P, I, I64, P], None),
            "orc_df_train_f64": ([P, P, P, P, P, I, I, P, I64, I, I, P, P], D),
            "orc_nrc_train_f64": ([P, P, P, I, P, I64, P], D),
            "orc_adam_ema_f64": ([I64, I64, D, D, D, D, D, D, I, P, P, P, P, P, P, P], None),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def det_expf(x):
    return np.float32(lib().orc_det_expf(float(x)))


def det_logf(x):
    return np.float32(lib().orc_det_logf(float(x)))


def half_round(x):
    return np.float32(lib().orc_half_round(float(x)))


def grid_index(L, inp, pos_dim=None):
    s = to_struct(L)
    inp = _c(inp, np.float32)
    n = inp.shape[0]
    g = L["grid"]
    out = np.zeros((n, g["levels"], 8), dtype=np.uint32)
    lib().orc_grid_index(ctypes.byref(s), _p(inp), n, inp.shape[1], g["dim_begin"] if pos_dim is None else pos_dim,
                         _p(out))
    return out


def encode_f64(L, params, inp):
    s = to_struct(L)
    params = _c(params, np.float64)
    inp = _c(inp, np.float64)
    out = np.zeros((inp.shape[0], L["enc_width"]), dtype=np.float64)
    lib().orc_encode_f64(ctypes.byref(s), _p(params), _p(inp), inp.shape[0], inp.shape[1], _p(out))
    return out


def mlp_f64(L, params, x):
    s = to_struct(L)
    params = _c(params, np.float64)
    x = _c(x, np.float64)
    out = np.zeros((x.shape[0], L["n_out_pad"]), dtype=np.float64)
    lib().orc_mlp_f64(ctypes.byref(s), _p(params), _p(x), x.shape[0], _p(out))
    return out


def head(logits, linear, u, ue, sample=None, x_given=None, fp32=False):
    """Softmax x N, inverse-CDF sample (where sample[i]), pdf.  Returns
    (x, pdf, bins, seg)."""
    dt = np.float32 if fp32 else np.float64
    logits = _c(logits, dt)
    n, N = logits.shape
    u, ue = _c(u, dt), _c(ue, dt)
    x = np.zeros(n, dt)
    pdf = np.zeros(n, dt)
    bins = np.zeros((n, N), dt)
    seg = np.zeros(n, np.int32)
    smp = None if sample is None else _c(sample, np.uint8)
    xg = None if x_given is None else _c(x_given, dt)
    fn = lib().orc_head_f32 if fp32 else lib().orc_head_f64
    fn(_p(logits), n, N, int(linear), _p(u), _p(ue), _p(smp), _p(xg), _p(x), _p(pdf), _p(bins), _p(seg))
    return x, pdf, bins, seg


def query_f64(Lt, Lp, pt, pp, feat, D, u, wi_in, linear, guided=None, with_logits=False):
    st, sp = to_struct(Lt), to_struct(Lp)
    pt, pp = _c(pt, np.float32), _c(pp, np.float32)
    feat, u, wi_in = _c(feat, np.float32), _c(u, np.float32), _c(wi_in, np.float32)
    n = feat.shape[0]
    out = np.zeros((n, 4), np.float32)
    lg = np.zeros((n, 128), np.float64) if with_logits else None
    g = None if guided is None else _c(guided, np.uint8)
    lib().orc_query_f64(ctypes.byref(st), ctypes.byref(sp), _p(pt), _p(pp), _p(feat), feat.shape[1], D, n, _p(u),
                        _p(g), _p(wi_in), int(linear), _p(out), _p(lg))
    return (out, lg) if with_logits else out


def query_f32_omp(Lt, Lp, pt, pp, feat, D, u, wi_in, linear, guided=None):
    """FP32 reference-order query on all host cores; returns (wi, threads)."""
    st, sp = to_struct(Lt), to_struct(Lp)
    pt, pp = _c(pt, np.float32), _c(pp, np.float32)
    feat, u, wi_in = _c(feat, np.float32), _c(u, np.float32), _c(wi_in, np.float32)
    n = feat.shape[0]
    out = np.zeros((n, 4), np.float32)
    g = None if guided is None else _c(guided, np.uint8)
    thr = lib().orc_query_f32_omp(ctypes.byref(st), ctypes.byref(sp), _p(pt), _p(pp), _p(feat), feat.shape[1], D, n,
                                  _p(u), _p(g), _p(wi_in), int(linear), _p(out))
    return out, thr


def nrc_f64(L, params, feat):
    s = to_struct(L)
    params, feat = _c(params, np.float32), _c(feat, np.float32)
    out = np.zeros((feat.shape[0], 3), np.float64)
    lib().orc_nrc_f64(ctypes.byref(s), _p(params), _p(feat), feat.shape[1], feat.shape[0], _p(out))
    return out


def nrc_f32_omp(L, params, feat):
    s = to_struct(L)
    params, feat = _c(params, np.float32), _c(feat, np.float32)
    out = np.zeros((feat.shape[0], 3), np.float32)
    lib().orc_nrc_f32_omp(ctypes.byref(s), _p(params), _p(feat), feat.shape[1], feat.shape[0], _p(out))
    return out


def df_train_f64(Lt, Lp, pt, pp, feat, D, target, linear, grads=True, half_bins=False):
    st, sp = to_struct(Lt), to_struct(Lp)
    pt, pp = _c(pt, np.float64), _c(pp, np.float64)
    feat, target = _c(feat, np.float32), _c(target, np.float32)
    gt = np.zeros(Lt["n_params"], np.float64) if grads else None
    gp = np.zeros(Lp["n_params"], np.float64) if grads else None
    loss = lib().orc_df_train_f64(ctypes.byref(st), ctypes.byref(sp), _p(pt), _p(pp), _p(feat), feat.shape[1], D,
                                  _p(target), feat.shape[0], int(linear), int(half_bins), _p(gt), _p(gp))
    return loss, gt, gp


def nrc_train_f64(L, params, feat, target, grads=True):
    s = to_struct(L)
    params, feat, target = _c(params, np.float64), _c(feat, np.float32), _c(target, np.float32)
    g = np.zeros(L["n_params"], np.float64) if grads else None
    loss = lib().orc_nrc_train_f64(ctypes.byref(s), _p(params), _p(feat), feat.shape[1], _p(target), feat.shape[0],
                                   _p(g))
    return loss, g


def adam_ema_f64(n_matrix, lr, b1, b2, eps, l2, decay, step, grad, master, m, v, steps, w_lowp=None, ema=None):
    grad, master = _c(grad, np.float64), _c(master, np.float64).copy()
    m, v, steps = _c(m, np.float64).copy(), _c(v, np.float64).copy(), _c(steps, np.uint32).copy()
    ema = None if ema is None else _c(ema, np.float64).copy()
    wl = None if w_lowp is None else _c(w_lowp, np.float64)
    lib().orc_adam_ema_f64(master.size, n_matrix, lr, b1, b2, eps, l2, decay, step, _p(grad), _p(master), _p(m),
                           _p(v), _p(steps), _p(wl), _p(ema))
    return master, m, v, steps, ema


# ----------------------------------------------------------------------------
# Training records -> training rows (generate_training_data_AR, train.cu:6-96;
# generate_training_data_nrc, train.cu:98-151), FP64.  Records are the SoA
# fields of GuidedPixelStateBuffer (workqueue.h:26-144) as [max_depth][n_pixels]
# arrays; rows come out pixel-major (depth inner) for the sampled training
# pixels offset + t * stride.
# ----------------------------------------------------------------------------
M_EPSILON = 1e-5


def spherical_to_cartesian(theta, phi):  # math_utils.h:198-208
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    return np.stack([st * cp, st * sp, ct], -1)


def uniform_sphere_to_square(v):  # math_utils.h:233-240
    vn = v / np.linalg.norm(v, axis=-1, keepdims=True)
    phi = np.arctan2(vn[..., 1], vn[..., 0]) / (2 * np.pi)
    return np.stack([(1.0 - vn[..., 2]) * 0.5, np.where(phi < 0, phi + 1.0, phi)], -1)


def square_to_uniform_sphere(u):  # math_utils.h:226-231
    z = 1.0 - 2.0 * u[..., 0]
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    phi = 2 * np.pi * u[..., 1]
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)


def normalize_spatial(p, aabb):  # train.h:272-276 (AABB inflated by 0.5 % of its diagonal)
    lo, hi = np.asarray(aabb[:3], np.float64), np.asarray(aabb[3:], np.float64)
    d = np.linalg.norm(hi - lo) * 0.005
    lo, hi = lo - d, hi + d
    return 0.5 + (p - (lo + hi) * 0.5) / (hi - lo)


def _vertices(rec, n_train, offset, stride, skip_last):
    """(depth, pixel) of every vertex considered, pixel-major with depth inner."""
    D = rec["pos"].shape[0]
    pix = offset + np.arange(n_train, dtype=np.int64) * stride
    depth = np.minimum(rec["cur_depth"][pix].astype(np.int64), D) - (1 if skip_last else 0)
    k = np.arange(D)[None, :].repeat(n_train, 0)
    p = pix[:, None].repeat(D, 1)
    m = k < depth[:, None]
    return k[m], p[m]


def _record_features(rec, k, p, aabb):
    pos = normalize_spatial(rec["pos"][k, p].astype(np.float64), aabb)
    wo = rec["wo"][k, p].astype(np.float64)
    dsh = (spherical_to_cartesian(wo[:, 0], wo[:, 1]) + 1.0) * 0.5  # warp_direction_for_sh, train.h:1243-1245
    nrm = rec["normal"][k, p].astype(np.float64)
    rough = 1.0 - np.exp(-rec["roughness"][k, p].astype(np.float64))  # warp_roughness_for_ob, train.h:1247-1249
    return np.concatenate([pos, dsh, (nrm[:, 0] / np.pi)[:, None], (nrm[:, 1] / (2 * np.pi))[:, None],
                           rough[:, None]], 1)


def _local_radiance(rec, k, p):
    L = rec["L"][k, p].astype(np.float64)
    thp = rec["thp"][k, p].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        L = np.where(thp > M_EPSILON, L / np.where(thp > M_EPSILON, thp, 1.0), 0.0)
    return L * rec["mis_weight"][k, p].astype(np.float64)[:, None]


def training_targets_f64(rec, n_train, offset, stride, aabb):
    """DF training rows: features [m][9] and targets [m][9] = (wi square 2,
    L 3, wiPdf, bsdfVal 3 left at 0 as in the reference), pixel-major."""
    k, p = _vertices(rec, n_train, offset, stride, False)
    keep = rec["delta"][k, p] == 0
    k, p = k[keep], p[keep]
    f = _record_features(rec, k, p, aabb)
    d = rec["dir"][k, p].astype(np.float64)
    wi = uniform_sphere_to_square(spherical_to_cartesian(d[:, 0], d[:, 1])).reshape(-1, 2)
    L = _local_radiance(rec, k, p)
    pdf = rec["wi_pdf"][k, p].astype(np.float64)
    # the reference tests L before the bsdfVal product (train.cu:80-84)
    ok = ~(np.isnan(f[:, :6]).any(1) | np.isnan(wi).any(1) | np.isnan(pdf) | (pdf == 0) | np.isnan(L).any(1))
    L = L * rec["bsdf_val"][k, p].astype(np.float64)
    t = np.concatenate([wi, L, pdf[:, None], np.zeros((len(pdf), 3))], 1)
    return f[ok].reshape(-1, 9), t[ok].reshape(-1, 9)


def nrc_training_targets_f64(rec, n_train, offset, stride, aabb):
    """NRC training rows (the last recorded depth has no radiance estimate and
    is skipped): features [m][15], targets [m][3] = L * bsdfVal / (wiPdf + eps)."""
    k, p = _vertices(rec, n_train, offset, stride, True)
    keep = (rec["delta"][k, p] == 0) & (rec["miss"][k, p] == 0)
    k, p = k[keep], p[keep]
    f = np.concatenate([_record_features(rec, k, p, aabb), rec["diffuse"][k, p].astype(np.float64),
                        rec["specular"][k, p].astype(np.float64)], 1)
    L = _local_radiance(rec, k, p) * rec["bsdf_val"][k, p].astype(np.float64)
    L = L / (rec["wi_pdf"][k, p].astype(np.float64)[:, None] + M_EPSILON)
    ok = ~(np.isnan(f[:, :6]).any(1) | np.isnan(L).any(1))
    return f[ok].reshape(-1, 15), L[ok].reshape(-1, 3)


def df_batches(n_samples, batch_size, batch_per_frame, min_batch=65536):
    """Batch schedule of LearnedDistribution::trainStep (learnedDistribution.cpp:226-236)."""
    out = []
    for it in range(min(n_samples // batch_size + 1, batch_per_frame)):
        b = min(n_samples - it * batch_size, batch_size)
        b -= b % 128
        if b < min_batch:
            break
        out.append((it * batch_size, b))
    return out, min(n_samples // batch_size + 1, batch_per_frame)


def nrc_batches(n_samples, batch_size, batch_per_frame, min_batch=65536):
    """Batch schedule of NRC::trainStep (nrc.cpp:127-138): the slice of batch
    `it` starts at it * (its own rounded size), as the reference indexes it."""
    n_b = max(n_samples // batch_size + 1, batch_per_frame)
    eff = min(n_samples // n_b, batch_size)
    out = []
    for it in range(n_b):
        b = min(n_samples - it * eff, eff)
        b -= b % 128
        if b < min_batch:
            break
        out.append((it * b, b))
    return out, n_b


def nrc_substitute_f64(rec, n_train, offset, stride, aabb, nrc_rgb):
    """The records after the radiance-cache substitution (OPTIMIZE_ON_NRC = 4:
    train.cu:156-213, 265-308; workqueue.h:113-138), given the cache's output
    nrc_rgb [m][3] for the inference inputs in the order this returns them.
    Returns (inputs [m][15], vertices [m][2] (depth, pixel), new L, new cur_depth)."""
    D = rec["pos"].shape[0]
    k, p = _vertices(rec, n_train, offset, stride, False)
    keep = (rec["delta"][k, p] == 0) & (rec["miss"][k, p] == 0)
    k, p = k[keep], p[keep]
    f = np.concatenate([_record_features(rec, k, p, aabb), rec["diffuse"][k, p].astype(np.float64),
                        rec["specular"][k, p].astype(np.float64)], 1)
    L = _local_radiance(rec, k, p) * rec["bsdf_val"][k, p].astype(np.float64)
    L = L / (rec["wi_pdf"][k, p].astype(np.float64)[:, None] + M_EPSILON)
    ok = ~(np.isnan(f[:, :6]).any(1) | np.isnan(L).any(1))
    k, p, f = k[ok], p[ok], f[ok]
    depth = rec["cur_depth"].astype(np.int64).copy()
    pix = offset + np.arange(n_train, dtype=np.int64) * stride
    depth[pix] = np.maximum(depth[pix] - 1, 0)
    newL = rec["L"].astype(np.float64).copy()
    Le = rec.get("Le")
    valid = lambda pp, d: 0 <= d < D and d < depth[pp]
    if nrc_rgb is None:
        return f, np.stack([k, p], 1), newL, depth
    for i in range(len(k)):   # replaceIndirectRadianceSkipDelta(p, k - 1, R)
        kk, pp = int(k[i]), int(p[i])
        d = kk - 1
        while d >= 0 and rec["delta"][d, pp]:
            d -= 1
        if not valid(pp, d):
            continue
        le = Le[kk - 1, pp].astype(np.float64) if Le is not None else 0.0
        newL[d, pp] = rec["thp"][kk - 1, pp].astype(np.float64) * (nrc_rgb[i] + le)
    for i in range(len(k)):   # divideRadiance(p, k, R)
        kk, pp = int(k[i]), int(p[i])
        if not valid(pp, kk):
            continue
        R = nrc_rgb[i]
        newL[kk, pp] = np.where(R > M_EPSILON, newL[kk, pp] / np.where(R > M_EPSILON, R, 1.0), newL[kk, pp])
    return f, np.stack([k, p], 1), newL, depth


def scatter_f64(f, wi, guided, mix, bsdf_pdf, bsdf_val, cos_theta, thp):
    """handleIntersections' scatter step (integrator.cpp:203-299) once the BSDF
    is evaluated: (thp', scatter pdf, MIS weight, alive)."""
    wi = np.asarray(wi, np.float64)
    pg = np.abs(wi[:, 2]) * np.abs(wi[:, 3]) * np.float64(np.float32(0.07957747154594766788))
    pb = np.asarray(bsdf_pdf, np.float64)
    g = np.zeros(len(pb), bool) if guided is None else np.asarray(guided) != 0
    m = np.zeros(len(pb), bool) if mix is None else np.asarray(mix) != 0
    p = np.where(g, (1 - f) * pg + f * pb, np.where(m, pb * f + (1 - f) * pg, pb))
    bv = np.asarray(bsdf_val, np.float64)
    alive = (p != 0) & (bv != 0).any(1)
    thp = np.asarray(thp, np.float64).copy()
    with np.errstate(divide="ignore", invalid="ignore"):
        upd = thp * bv * np.abs(np.asarray(cos_theta, np.float64))[:, None] / (p[:, None] + M_EPSILON)
        mis = (1 - f) * pg / p
    thp = np.where(alive[:, None], upd, thp)
    return thp, p, mis, alive


# ----------------------------------------------------------------------------
# The consumer stage around the query (GuidedPathTracer, integrator.cpp):
# sampleWi's decision and the record producers of GuidedPixelStateBuffer
# (workqueue.h:26-92), FP64, one row at a time.
# ----------------------------------------------------------------------------
def _normalized(v):
    """Eigen's normalized(): v / |v|, v itself when |v| = 0."""
    v = np.asarray(v, np.float64)
    n = math.sqrt(float(v @ v))
    return v / n if n > 0 else v


def cartesian_to_spherical(v):  # math_utils.h:183-190
    vn = _normalized(v)
    th = math.acos(min(max(vn[2], -1.0), 1.0))
    ph = math.atan2(vn[1], vn[0])
    return np.array([th, ph + 2 * math.pi if ph < 0 else ph])


def uniform_sphere_to_square_one(v):  # math_utils.h:233-240
    vn = _normalized(v)
    ph = math.atan2(vn[1], vn[0]) / (2 * math.pi)
    return np.array([(1.0 - vn[2]) * 0.5, ph + 1.0 if ph < 0 else ph])


def sample_wi_f64(rr_prob, f, guiding_enabled, max_guided_depth, depth, smooth, u_rr, u_mix, bsdf_wi, bsdf_pdf,
                  bsdf_f, wi, thp=None):
    """sampleWi (integrator.cpp:355-401) and the thp /= probRR of
    handleIntersections (integrator.cpp:211-212).  Returns (rr_killed, guided,
    mix, wi', thp')."""
    n = len(depth)
    killed, guided, mix = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    wi = np.asarray(wi, np.float64).copy()
    thp = None if thp is None else np.asarray(thp, np.float64).copy()
    for i in range(n):
        if u_rr[i] >= rr_prob:
            killed[i] = 1
            continue
        if thp is not None:
            thp[i] /= rr_prob
        on = bool(guiding_enabled) and depth[i] < max_guided_depth and smooth[i] != 0
        g = on and (f == 0 or u_mix[i] >= f)
        guided[i] = g
        mix[i] = (not g) and on and f < 1          # integrator.cpp:256-261
        wi[i, 2] = wi[i, 3] = 1.0
        if not g and bsdf_pdf[i] != 0 and np.any(np.asarray(bsdf_f[i]) != 0):
            wi[i, :2] = uniform_sphere_to_square_one(bsdf_wi[i])
    return killed, guided, mix, wi, thp


def _is_training_pixel(p, offset, stride):  # guided.h:26-28, uint32 arithmetic
    return ((int(p) - int(offset)) & 0xffffffff) % int(stride) == 0


def _last_bounce(rec, p, miss, sd=None, i=None):
    """incrementDepthLastBounceNRC (workqueue.h:60-81); sd None: an empty
    ShadingData (zero position, reflectances and directions, roughness 1,
    diffuse BSDF)."""
    d = int(rec["cur_depth"][p])
    if d >= rec["pos"].shape[0]:
        return
    rec["cur_depth"][p] = d + 1
    if sd is None:
        rec["pos"][d, p] = 0.0
        rec["normal"][d, p] = cartesian_to_spherical(np.zeros(3))
        rec["wo"][d, p] = cartesian_to_spherical(np.zeros(3))
        rec["roughness"][d, p] = 1.0
        rec["diffuse"][d, p] = rec["specular"][d, p] = 0.0
        rec["delta"][d, p] = 0
    else:
        rec["pos"][d, p] = sd["pos"][i]
        rec["normal"][d, p] = cartesian_to_spherical(sd["normal"][i])
        rec["wo"][d, p] = cartesian_to_spherical(sd["wo"][i])
        rec["roughness"][d, p] = sd["roughness"][i]
        rec["diffuse"][d, p] = sd["diffuse"][i]
        rec["specular"][d, p] = sd["specular"][i]
        rec["delta"][d, p] = 0 if sd["smooth"][i] else 1
    rec["miss"][d, p] = 1 if miss else 0
    rec["L"][d, p] = rec["thp"][d, p] = 0.0
    if "Le" in rec:
        rec["Le"][d, p] = 0.0


def record_scatter_f64(rec, sd, pixel, depth, train_enabled, max_train_depth, offset, stride, mis_aware,
                       product_with_cosine, rr_killed=None, alive=None, wi_dir=None, thp=None, scatter_pdf=None,
                       mis=None, bsdf_val=None, cos_theta=None, delta=None):
    """The record step of handleIntersections (integrator.cpp:277-296):
    incrementDepth (workqueue.h:26-58) for scattered rows,
    incrementDepthLastBounceNRC for the others (and handleNRCLastBounce,
    integrator.cpp:193-201, when alive is None).  Updates rec (numpy arrays
    [max_depth][n_pixels](k), cur_depth [n_pixels]) in place."""
    D, N = rec["pos"].shape[0], rec["pos"].shape[1]
    for i in range(len(pixel)):
        p = int(pixel[i])
        if rr_killed is not None and rr_killed[i]:
            continue
        if p < 0 or p >= N:
            continue
        if not (train_enabled and depth[i] < max_train_depth and _is_training_pixel(p, offset, stride)):
            continue
        if alive is None or not alive[i]:
            _last_bounce(rec, p, False, sd, i)
            continue
        d = int(rec["cur_depth"][p])
        if d >= D:
            continue
        rec["L"][d, p] = 0.0
        rec["thp"][d, p] = thp[i]
        rec["pos"][d, p] = sd["pos"][i]
        rec["dir"][d, p] = cartesian_to_spherical(wi_dir[i])
        rec["wi_pdf"][d, p] = scatter_pdf[i]
        rec["mis_weight"][d, p] = mis[i] if mis_aware else 1.0
        rec["delta"][d, p] = 1 if (delta is not None and delta[i]) else 0
        rec["miss"][d, p] = 0
        rec["normal"][d, p] = cartesian_to_spherical(sd["normal"][i])
        rec["roughness"][d, p] = sd["roughness"][i]
        rec["diffuse"][d, p] = sd["diffuse"][i]
        rec["specular"][d, p] = sd["specular"][i]
        b = np.asarray(bsdf_val[i], np.float64)
        rec["bsdf_val"][d, p] = b * abs(float(cos_theta[i])) if product_with_cosine else b
        rec["wo"][d, p] = cartesian_to_spherical(sd["wo"][i])
        if "Le" in rec:
            rec["Le"][d, p] = 0.0
        rec["cur_depth"][p] = d + 1


def record_radiance_f64(rec, pixel, contrib, train_enabled, max_train_depth, offset, stride, Le=None, miss=None,
                        depth=None):
    """handleEmissiveHit / handleMiss (integrator.cpp:144-191): recordRadiance
    (workqueue.h:83-92), overwriteLe (workqueue.h:98-106) and, for misses at
    depth <= max_train_depth, the placeholder vertex.  In place."""
    D, N = rec["pos"].shape[0], rec["pos"].shape[1]
    for i in range(len(pixel)):
        p = int(pixel[i])
        if p < 0 or p >= N or not (train_enabled and _is_training_pixel(p, offset, stride)):
            continue
        cd = int(rec["cur_depth"][p])
        for d in range(min(cd, D)):
            rec["L"][d, p] = rec["L"][d, p] + contrib[i]
        last = cd - 1
        if Le is not None and "Le" in rec and 0 <= last < D:
            rec["Le"][last, p] = Le[i]
        if miss is not None and miss[i] and depth[i] <= max_train_depth:
            _last_bounce(rec, p, True)
