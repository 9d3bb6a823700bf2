"""Python mirror of the guide's operator interface over the C ABI (libdfg.so).

The reference drives the guide through the C++ class LearnedDistribution
(src/render/guided/learnedDistribution.h:37-168) and NRC (nrc.h:30-97); this
module exposes the same operations — reset/create, query (inferenceStep),
pdf (AutoRegressive::pdf), train step, NRC infer/train, save/load — on device
tensors (torch is only used for device memory and streams).  Every call goes
through the HIP library; there is no CPU fallback: loading fails loudly when
libdfg.so is missing.
"""
import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdfg.so")

DFG_NET_THETA, DFG_NET_PHI, DFG_NET_NRC = 0, 1, 2
DFG_PARAMS_MASTER, DFG_PARAMS_TRAIN, DFG_PARAMS_INFERENCE = 0, 1, 2

STATUS = {
    0: "DFG_OK",
    1: "DFG_ERR_INVALID_ARG",
    2: "DFG_ERR_BAD_SIZE",
    3: "DFG_ERR_CONFIG",
    4: "DFG_ERR_HIP",
    5: "DFG_ERR_OOM",
    6: "DFG_ERR_NOT_INIT",
    7: "DFG_ERR_IO",
    8: "DFG_ERR_OVER_BUDGET",
}

# Every symbol include/dfg.h declares.
EXPORTS = [
    "dfg_create", "dfg_destroy", "dfg_last_error", "dfg_get_info", "dfg_get_params", "dfg_get_grad",
    "dfg_set_params", "dfg_save", "dfg_load", "dfg_query", "dfg_query_noalloc", "dfg_query_plain", "dfg_pdf",
    "dfg_train_step", "dfg_nrc_infer", "dfg_nrc_train_step", "dfg_train_step_plain",
    "dfg_nrc_train_step_plain", "dfg_training_targets", "dfg_nrc_training_targets", "dfg_train_records",
    "dfg_nrc_train_records", "dfg_nrc_substitute_radiance", "dfg_pdf_image", "dfg_directions", "dfg_scatter", "dfg_sample_wi",
    "dfg_record_scatter", "dfg_record_radiance", "dfg_prepare_features",
    "dfg_df_head", "dfg_logits", "dfg_encode", "dfg_grid_index",
    "dfg_wavefront_create", "dfg_wavefront_destroy", "dfg_wavefront_set_tuning", "dfg_wavefront_get_tuning",
    "dfg_wavefront_run", "dfg_wavefront_autotune",
    "dfg_server_create", "dfg_server_destroy", "dfg_server_submit", "dfg_server_submit_train",
    "dfg_server_progress", "dfg_server_get_stats", "dfg_server_drain",
]


class DfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class DfgInfo(ctypes.Structure):
    _fields_ = [
        ("n_bins_theta", ctypes.c_int), ("n_bins_phi", ctypes.c_int),
        ("interpolation", ctypes.c_int), ("n_input_dims", ctypes.c_int),
        ("has_nrc", ctypes.c_int), ("enc_width", ctypes.c_int * 3),
        ("n_hidden", ctypes.c_int * 3), ("n_out_padded", ctypes.c_int * 3),
        ("n_params", ctypes.c_size_t * 3), ("n_matrix_params", ctypes.c_size_t * 3),
        ("max_batch", ctypes.c_int64),
    ]


class DfgWavefrontTuning(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int), ("tiles_per_fetch", ctypes.c_int),
                ("grid_prepass", ctypes.c_int), ("ms", ctypes.c_float)]


class DfgServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_int64), ("max_batch_requests", ctypes.c_int),
                ("n_streams", ctypes.c_int), ("train_every", ctypes.c_int), ("round_rows", ctypes.c_int64)]


class DfgRequest(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("n", ctypes.c_int64), ("feat", ctypes.c_void_p),
                ("feat_stride", ctypes.c_int), ("u", ctypes.c_void_p), ("guided", ctypes.c_void_p),
                ("wi", ctypes.c_void_p), ("nrc_rgb", ctypes.c_void_p)]


class DfgServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("train_steps", ctypes.c_int64), ("rows", ctypes.c_int64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("busy_s", ctypes.c_double), ("peak_rows", ctypes.c_int64)]


RECORD_FIELDS = ("cur_depth", "pos", "dir", "wo", "normal", "roughness", "L", "thp", "bsdf_val", "diffuse",
                 "specular", "wi_pdf", "mis_weight", "delta", "miss", "Le")


class DfgRecords(ctypes.Structure):
    _fields_ = [("max_depth", ctypes.c_int), ("n_pixels", ctypes.c_int64)] + \
               [(f, ctypes.c_void_p) for f in RECORD_FIELDS]


SHADING_FIELDS = ("pos", "normal", "wo", "roughness", "diffuse", "specular", "smooth")


class DfgShading(ctypes.Structure):
    _fields_ = [(f, ctypes.c_void_p) for f in SHADING_FIELDS]


def shading_struct(sd):
    """dfg_shading view of a dict of device tensors [n(, k)]."""
    r = DfgShading()
    for f in SHADING_FIELDS:
        t = sd.get(f)
        setattr(r, f, None if t is None else t.data_ptr())
    return r


def records_struct(rec):
    """dfg_records view of a dict of device tensors shaped [max_depth, n_pixels(, k)]
    (cur_depth: [n_pixels] uint32 / int32)."""
    d, n = rec["pos"].shape[0], rec["pos"].shape[1]
    r = DfgRecords()
    r.max_depth, r.n_pixels = d, n
    for f in RECORD_FIELDS:
        t = rec.get(f)
        setattr(r, f, None if t is None else t.data_ptr())
    return r


_lib = None


def load_library(path=LIB_PATH):
    """Load libdfg.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            "libdfg.so not built (%s); run python guiding-df_amd/build.py" % path)
    lib = ctypes.CDLL(path)
    P, I, I64, SZ, U32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t, ctypes.c_uint32
    sig = {
        "dfg_create": ([ctypes.c_char_p, ctypes.POINTER(P)], I),
        "dfg_destroy": ([P], I),
        "dfg_last_error": ([P], ctypes.c_char_p),
        "dfg_get_info": ([P, ctypes.POINTER(DfgInfo)], I),
        "dfg_get_params": ([P, I, I, P, SZ], I),
        "dfg_set_params": ([P, I, P, SZ], I),
        "dfg_get_grad": ([P, I, P, SZ], I),
        "dfg_save": ([P, ctypes.c_char_p], I),
        "dfg_load": ([P, ctypes.c_char_p], I),
        "dfg_query": ([P, P, I64, P, I, P, P, P, P], I),
        "dfg_query_noalloc": ([P, P, I64, P, I, P, P, P, P], I),
        "dfg_query_plain": ([P, P, I64, P, I, P, P, P, P], I),
        "dfg_pdf": ([P, P, I64, P, I, P, P], I),
        "dfg_train_step": ([P, P, I64, P, I, P, P], I),
        "dfg_train_step_plain": ([P, P, I64, P, I, P, P], I),
        "dfg_nrc_infer": ([P, P, I64, P, I, P], I),
        "dfg_nrc_train_step": ([P, P, I64, P, I, P, P], I),
        "dfg_nrc_train_step_plain": ([P, P, I64, P, I, P, P], I),
        "dfg_training_targets": ([P, P, ctypes.POINTER(DfgRecords), I64, U32, U32, P, P, P, I64, P], I),
        "dfg_nrc_training_targets": ([P, P, ctypes.POINTER(DfgRecords), I64, U32, U32, P, P, P, I64, P], I),
        "dfg_train_records": ([P, P, ctypes.POINTER(DfgRecords), I64, U32, U32, P, I64, I,
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_float)], I),
        "dfg_nrc_train_records": ([P, P, ctypes.POINTER(DfgRecords), I64, U32, U32, P, I64, I,
                                   ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_float)], I),
        "dfg_nrc_substitute_radiance": ([P, P, ctypes.POINTER(DfgRecords), I64, U32, U32, P,
                                         ctypes.POINTER(ctypes.c_int64)], I),
        "dfg_pdf_image": ([P, P, P, I, I, P], I),
        "dfg_directions": ([P, I64, P, P, P], I),
        "dfg_scatter": ([P, I64, ctypes.c_float, P, P, P, P, P, P, P, P, P, P], I),
        "dfg_sample_wi": ([P, I64, ctypes.c_float, ctypes.c_float, I, I] + [P] * 12, I),
        "dfg_record_scatter": ([P, I64, P, P, I, I, ctypes.c_uint32, ctypes.c_uint32, I, I] + [P] * 11, I),
        "dfg_record_radiance": ([P, I64, P, I, I, ctypes.c_uint32, ctypes.c_uint32] + [P] * 5, I),
        "dfg_prepare_features": ([P, I64, P, P, P, P, P, P], I),
        "dfg_df_head": ([P, I64, I, I, I, P, P, P, P, P, P, P, P], I),
        "dfg_logits": ([P, P, I64, P, I, P, P, P], I),
        "dfg_encode": ([P, P, I, I64, P, I, P], I),
        "dfg_grid_index": ([P, P, I, I64, P, I, P], I),
        "dfg_wavefront_create": ([P, I, I64, ctypes.POINTER(P)], I),
        "dfg_wavefront_destroy": ([P], I),
        "dfg_wavefront_set_tuning": ([P, ctypes.POINTER(DfgWavefrontTuning)], I),
        "dfg_wavefront_get_tuning": ([P, ctypes.POINTER(DfgWavefrontTuning)], I),
        "dfg_wavefront_run": ([P, P, I, P, P, I, P, P, P], I),
        "dfg_wavefront_autotune": ([P, P, I, P, P, I, P, P, P, ctypes.POINTER(DfgWavefrontTuning)], I),
        "dfg_server_create": ([P, ctypes.POINTER(DfgServerConfig), ctypes.POINTER(P)], I),
        "dfg_server_destroy": ([P], I),
        "dfg_server_submit": ([P, ctypes.POINTER(DfgRequest)], I),
        "dfg_server_submit_train": ([P, I64, P, I, P, P, P], I),
        "dfg_server_progress": ([P, I], I),
        "dfg_server_get_stats": ([P, ctypes.POINTER(DfgServerStats)], I),
        "dfg_server_drain": ([P, P, P, P, I, ctypes.POINTER(I)], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _check(lib, ctx, code):
    if code != 0:
        msg = lib.dfg_last_error(ctx)
        raise DfgError(code, msg.decode() if msg else "")


# ----------------------------------------------------------------------------
# Configurations.  reference_config() is the reference's own pair of files
# (common/configs/nn/base_product_df.json and base_hashgrid_nrc_df.json).
# ----------------------------------------------------------------------------
def _ema(decay, lr, eps, beta2=0.99):
    return {"otype": "EMA", "decay": decay, "nested": {
        "otype": "Adam", "learning_rate": lr, "beta1": 0.9, "beta2": beta2,
        "epsilon": eps, "l2_reg": 1e-6}}


def _nrc_section():
    return {
        "loss": {"otype": "RelativeL2Luminance"},
        "optimizer": _ema(0.99, 8e-3, 1e-8),
        "encoding": {"otype": "Composite", "nested": [
            {"n_dims_to_encode": 3, "otype": "Frequency", "n_frequencies": 12},
            {"otype": "OneBlob", "n_bins": 4, "n_dims_to_encode": 6},
            {"n_dims_to_encode": 6, "otype": "Identity"}]},
        "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                    "n_neurons": 64, "n_hidden_layers": 5},
    }


def reference_config(with_nrc=True, seed=7272):
    """base_product_df.json (DenseGrid 8x4 smoothstep + SH4 + OneBlob4, phi adds
    TriangleWave 12) with the reference's compile-time bins 16/32."""
    grid = {"otype": "DenseGrid", "n_levels": 8, "n_features_per_level": 4, "base_resolution": 8,
            "per_level_scale": 1.405, "interpolation": "Smoothstep", "n_dims_to_encode": 3}
    th = {"otype": "Composite", "nested": [
        grid, {"n_dims_to_encode": 3, "otype": "SphericalHarmonics", "degree": 4},
        {"otype": "OneBlob", "n_bins": 4, "n_dims_to_encode": 3}]}
    ph = {"otype": "Composite", "nested": th["nested"] + [
        {"n_dims_to_encode": 1, "otype": "TriangleWave", "n_frequencies": 12}]}
    cfg = {
        "dfg": {"n_bins_theta": 16, "n_bins_phi": 32, "interpolation": "linear", "seed": seed,
                "n_input_dims": 9},
        "nn": {"loss": {"otype": "L2"}, "optimizer": _ema(0.95, 3e-2, 1e-15),
               "encoding": th, "phi_encoding": ph,
               "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                           "n_neurons": 64, "n_hidden_layers": 3}},
    }
    if with_nrc:
        cfg["nrc"] = _nrc_section()
    return cfg


def plain_head_config(n_bins=32, seed=7272, interpolation="linear"):
    """Config 1: 8-d features, identity encoding; phi adds OneBlob(16) of theta."""
    net = {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
           "n_neurons": 64, "n_hidden_layers": 3}
    return {
        "dfg": {"n_bins_theta": n_bins, "n_bins_phi": n_bins, "interpolation": interpolation,
                "seed": seed, "n_input_dims": 8},
        "nn": {"optimizer": _ema(0.95, 1e-2, 1e-15),
               "encoding": {"otype": "Identity"},
               "phi_encoding": {"otype": "Composite", "nested": [
                   {"otype": "Identity", "n_dims_to_encode": 8},
                   {"otype": "OneBlob", "n_bins": 16, "n_dims_to_encode": 1}]},
               "network": net},
    }


def hash_config(n_bins_theta=32, n_bins_phi=32, with_nrc=True, seed=7272, n_levels=16,
                log2_hashmap_size=19, lr=1e-2):
    """Configs 2-4: 16-level hash grid (F=2, T=2^19, base 16, scale 2) + SH4 +
    OneBlob4 per net; phi adds TriangleWave 12; 64 x 3 MLPs."""
    grid = {"otype": "HashGrid", "n_levels": n_levels, "n_features_per_level": 2,
            "log2_hashmap_size": log2_hashmap_size, "base_resolution": 16, "per_level_scale": 2.0,
            "interpolation": "Linear", "n_dims_to_encode": 3}
    th = {"otype": "Composite", "nested": [
        grid, {"n_dims_to_encode": 3, "otype": "SphericalHarmonics", "degree": 4},
        {"otype": "OneBlob", "n_bins": 4, "n_dims_to_encode": 3}]}
    ph = {"otype": "Composite", "nested": th["nested"] + [
        {"n_dims_to_encode": 1, "otype": "TriangleWave", "n_frequencies": 12}]}
    cfg = {
        "dfg": {"n_bins_theta": n_bins_theta, "n_bins_phi": n_bins_phi, "interpolation": "linear",
                "seed": seed, "n_input_dims": 9},
        "nn": {"optimizer": _ema(0.95, lr, 1e-15), "encoding": th, "phi_encoding": ph,
               "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None",
                           "n_neurons": 64, "n_hidden_layers": 3}},
    }
    if with_nrc:
        cfg["nrc"] = _nrc_section()
    return cfg


class GuidingField:
    """Owner of one dfg context: the theta/phi nets (+ optional NRC)."""

    def __init__(self, config):
        self.lib = load_library()
        if isinstance(config, dict):
            config = json.dumps(config)
        ctx = ctypes.c_void_p()
        code = self.lib.dfg_create(config.encode(), ctypes.byref(ctx))
        if code != 0:
            raise DfgError(code, self.lib.dfg_last_error(None).decode())
        self.ctx = ctx
        self.info = DfgInfo()
        _check(self.lib, self.ctx, self.lib.dfg_get_info(self.ctx, ctypes.byref(self.info)))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.dfg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, code):
        _check(self.lib, self.ctx, code)

    # --- parameters ---------------------------------------------------------
    def n_params(self, net):
        return int(self.info.n_params[net])

    def get_params(self, net, which=DFG_PARAMS_MASTER):
        import numpy as np
        n = self.n_params(net)
        out = np.empty(n, dtype=np.float32)
        self._c(self.lib.dfg_get_params(self.ctx, net, which, out.ctypes.data_as(ctypes.c_void_p), n))
        return out

    def get_grad(self, net):
        """Last step's gradient, loss-scaled by 1024."""
        import numpy as np
        n = self.n_params(net)
        out = np.empty(n, dtype=np.float32)
        self._c(self.lib.dfg_get_grad(self.ctx, net, out.ctypes.data_as(ctypes.c_void_p), n))
        return out

    def set_params(self, net, params):
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._c(self.lib.dfg_set_params(self.ctx, net, p.ctypes.data_as(ctypes.c_void_p), p.size))

    def save(self, prefix):
        self._c(self.lib.dfg_save(self.ctx, prefix.encode()))

    def load(self, prefix):
        self._c(self.lib.dfg_load(self.ctx, prefix.encode()))

    # --- operators (device tensors) -----------------------------------------
    def query(self, feat, u, wi, guided=None, nrc_rgb=None, stream=None):
        n = feat.shape[0]
        self._c(self.lib.dfg_query(self.ctx, _stream(stream), n, _ptr(feat), feat.shape[1], _ptr(u),
                                   _ptr(guided), _ptr(wi), _ptr(nrc_rgb)))

    def query_noalloc(self, feat, u, wi, guided=None, nrc_rgb=None, stream=None):
        """dfg_query without device allocation (rows limited by the library's
        fixed scratch array)."""
        n = feat.shape[0]
        self._c(self.lib.dfg_query_noalloc(self.ctx, _stream(stream), n, _ptr(feat), feat.shape[1], _ptr(u),
                                           _ptr(guided), _ptr(wi), _ptr(nrc_rgb)))

    def query_plain(self, feat, u, wi, guided=None, nrc_rgb=None, stream=None):
        n = feat.shape[0]
        self._c(self.lib.dfg_query_plain(self.ctx, _stream(stream), n, _ptr(feat), feat.shape[1], _ptr(u),
                                         _ptr(guided), _ptr(wi), _ptr(nrc_rgb)))

    def pdf(self, feat, wi2, out, stream=None):
        self._c(self.lib.dfg_pdf(self.ctx, _stream(stream), feat.shape[0], _ptr(feat), feat.shape[1],
                                 _ptr(wi2), _ptr(out)))

    def train_step(self, feat, target, loss_out=None, stream=None):
        self._c(self.lib.dfg_train_step(self.ctx, _stream(stream), feat.shape[0], _ptr(feat), feat.shape[1],
                                        _ptr(target), _ptr(loss_out)))

    def train_step_plain(self, feat, target, loss_out=None, stream=None):
        """dfg_train_step on the plain FP32 per-batch path (the comparator)."""
        self._c(self.lib.dfg_train_step_plain(self.ctx, _stream(stream), feat.shape[0], _ptr(feat), feat.shape[1],
                                              _ptr(target), _ptr(loss_out)))

    def nrc_train_step_plain(self, feat, target_rgb, loss_out=None, stream=None):
        """dfg_nrc_train_step on the plain FP32 per-batch path (the comparator)."""
        self._c(self.lib.dfg_nrc_train_step_plain(self.ctx, _stream(stream), feat.shape[0], _ptr(feat),
                                                  feat.shape[1], _ptr(target_rgb), _ptr(loss_out)))

    # --- training from recorded path vertices ----------------------------------
    def training_targets(self, rec, n_train, aabb, feat, target, count, offset=0, stride=1, stream=None):
        r = records_struct(rec)
        self._c(self.lib.dfg_training_targets(self.ctx, _stream(stream), ctypes.byref(r), n_train, offset, stride,
                                              _ptr(aabb), _ptr(feat), _ptr(target), feat.shape[0], _ptr(count)))

    def nrc_training_targets(self, rec, n_train, aabb, feat, target_rgb, count, offset=0, stride=1, stream=None):
        r = records_struct(rec)
        self._c(self.lib.dfg_nrc_training_targets(self.ctx, _stream(stream), ctypes.byref(r), n_train, offset,
                                                  stride, _ptr(aabb), _ptr(feat), _ptr(target_rgb), feat.shape[0],
                                                  _ptr(count)))

    def _train_records(self, fn, rec, n_train, aabb, batch_size, batch_per_frame, offset, stride, stream):
        r = records_struct(rec)
        rows, steps, loss = ctypes.c_int64(0), ctypes.c_int(0), ctypes.c_float(0.0)
        self._c(fn(self.ctx, _stream(stream), ctypes.byref(r), n_train, offset, stride, _ptr(aabb), batch_size,
                   batch_per_frame, ctypes.byref(rows), ctypes.byref(steps), ctypes.byref(loss)))
        return rows.value, steps.value, loss.value

    def pdf_image(self, feat, image, stream=None):
        """Solid-angle pdf over the (u0, u1) pixel-centre grid of image [res_theta, res_phi]."""
        self._c(self.lib.dfg_pdf_image(self.ctx, _stream(stream), _ptr(feat), image.shape[0], image.shape[1],
                                       _ptr(image)))

    def nrc_substitute_radiance(self, rec, n_train, aabb, offset=0, stride=1, stream=None):
        """Replaces the records' indirect radiance by the radiance cache (writes
        rec["L"] and rec["cur_depth"]); returns the number of NRC queries."""
        r = records_struct(rec)
        m = ctypes.c_int64(0)
        self._c(self.lib.dfg_nrc_substitute_radiance(self.ctx, _stream(stream), ctypes.byref(r), n_train, offset,
                                                     stride, _ptr(aabb), ctypes.byref(m)))
        return m.value

    def train_records(self, rec, n_train, aabb, batch_size=524288, batch_per_frame=5, offset=0, stride=1,
                      stream=None):
        """LearnedDistribution::trainStep from records: (rows, steps run, loss)."""
        return self._train_records(self.lib.dfg_train_records, rec, n_train, aabb, batch_size, batch_per_frame,
                                   offset, stride, stream)

    def nrc_train_records(self, rec, n_train, aabb, batch_size=1048576, batch_per_frame=4, offset=0, stride=1,
                          stream=None):
        """NRC::trainStep from records: (rows, steps run, loss)."""
        return self._train_records(self.lib.dfg_nrc_train_records, rec, n_train, aabb, batch_size,
                                   batch_per_frame, offset, stride, stream)

    def nrc_infer(self, feat, rgb, stream=None):
        self._c(self.lib.dfg_nrc_infer(self.ctx, _stream(stream), feat.shape[0], _ptr(feat), feat.shape[1],
                                       _ptr(rgb)))

    def nrc_train_step(self, feat, target_rgb, loss_out=None, stream=None):
        self._c(self.lib.dfg_nrc_train_step(self.ctx, _stream(stream), feat.shape[0], _ptr(feat),
                                            feat.shape[1], _ptr(target_rgb), _ptr(loss_out)))

    def logits(self, feat, theta_in, lt, lp, stream=None):
        self._c(self.lib.dfg_logits(self.ctx, _stream(stream), feat.shape[0], _ptr(feat), feat.shape[1],
                                    _ptr(theta_in), _ptr(lt), _ptr(lp)))

    def encode(self, net, inp, out, stream=None):
        self._c(self.lib.dfg_encode(self.ctx, _stream(stream), net, inp.shape[0], _ptr(inp), inp.shape[1],
                                    _ptr(out)))

    def grid_index(self, net, inp, idx, stream=None):
        self._c(self.lib.dfg_grid_index(self.ctx, _stream(stream), net, inp.shape[0], _ptr(inp), inp.shape[1],
                                        _ptr(idx)))


class Wavefront:
    """All bounces of a wavefront frame in one persistent, graph-captured launch.
    Bounce batches are packed back to back in feat/u/guided/wi; `sizes` is a
    device int64 tensor with one entry per bounce."""

    def __init__(self, field, max_bounces, capacity):
        self.field = field
        self.lib = field.lib
        h = ctypes.c_void_p()
        field._c(self.lib.dfg_wavefront_create(field.ctx, max_bounces, capacity, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.dfg_wavefront_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, sizes, feat, u, wi, guided=None, stream=None):
        self.field._c(self.lib.dfg_wavefront_run(self.h, _stream(stream), sizes.numel(), _ptr(sizes), _ptr(feat),
                                                 feat.shape[1], _ptr(u), _ptr(guided), _ptr(wi)))

    def autotune(self, sizes, feat, u, wi, guided=None, stream=None):
        t = DfgWavefrontTuning()
        self.field._c(self.lib.dfg_wavefront_autotune(self.h, _stream(stream), sizes.numel(), _ptr(sizes), _ptr(feat),
                                                      feat.shape[1], _ptr(u), _ptr(guided), _ptr(wi),
                                                      ctypes.byref(t)))
        return {"blocks_per_cu": t.blocks_per_cu, "tiles_per_fetch": t.tiles_per_fetch,
                "grid_prepass": t.grid_prepass, "ms": t.ms}

    def set_tuning(self, blocks_per_cu, tiles_per_fetch, grid_prepass):
        t = DfgWavefrontTuning(blocks_per_cu, tiles_per_fetch, grid_prepass, 0.0)
        self.field._c(self.lib.dfg_wavefront_set_tuning(self.h, ctypes.byref(t)))


class Server:
    """Continuous-batching request server over one GuidingField (admission on
    the device; round_rows 0 picks the library's default)."""

    def __init__(self, field, budget_bytes=16 << 30, max_batch_requests=64, n_streams=2, train_every=8,
                 round_rows=0):
        self.field = field
        self.lib = field.lib
        cfg = DfgServerConfig(budget_bytes, max_batch_requests, n_streams, train_every, round_rows)
        h = ctypes.c_void_p()
        field._c(self.lib.dfg_server_create(field.ctx, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.dfg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, rid, feat, u, wi, guided=None, nrc_rgb=None):
        r = DfgRequest(rid, feat.shape[0], feat.data_ptr(), feat.shape[1], u.data_ptr(),
                       guided.data_ptr() if guided is not None else None, wi.data_ptr(),
                       nrc_rgb.data_ptr() if nrc_rgb is not None else None)
        self._keep[rid] = (feat, u, wi, guided, nrc_rgb)
        self.field._c(self.lib.dfg_server_submit(self.h, ctypes.byref(r)))

    def submit_train(self, feat, target, nrc_feat=None, nrc_target=None):
        self._keep[("train", len(self._keep))] = (feat, target, nrc_feat, nrc_target)
        self.field._c(self.lib.dfg_server_submit_train(self.h, feat.shape[0], _ptr(feat), feat.shape[1], _ptr(target),
                                                       _ptr(nrc_feat), _ptr(nrc_target)))

    def progress(self, wait=False):
        self.field._c(self.lib.dfg_server_progress(self.h, 1 if wait else 0))

    def stats(self):
        st = DfgServerStats()
        self.field._c(self.lib.dfg_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in DfgServerStats._fields_}

    def drain(self, max_records=1 << 16):
        import numpy as np
        ids = np.zeros(max_records, np.int64)
        lat = np.zeros(max_records, np.float64)
        ver = np.zeros(max_records, np.int64)
        cnt = ctypes.c_int(0)
        self.field._c(self.lib.dfg_server_drain(self.h, ids.ctypes.data_as(ctypes.c_void_p),
                                                lat.ctypes.data_as(ctypes.c_void_p),
                                                ver.ctypes.data_as(ctypes.c_void_p), max_records, ctypes.byref(cnt)))
        k = cnt.value
        return ids[:k], lat[:k], ver[:k]


def df_head(logits_theta, logits_phi, u, wi, guided=None, interpolation=1, bins_theta=None,
            bins_phi=None, seg=None, stream=None):
    lib = load_library()
    n = logits_theta.shape[0]
    code = lib.dfg_df_head(_stream(stream), n, logits_theta.shape[1], logits_phi.shape[1], interpolation,
                           _ptr(logits_theta), _ptr(logits_phi), _ptr(u), _ptr(guided), _ptr(wi),
                           _ptr(bins_theta), _ptr(bins_phi), _ptr(seg))
    if code != 0:
        raise DfgError(code, "dfg_df_head")


def directions(wi, dir=None, pdf=None, stream=None):
    """Directions (squareToUniformSphere) and solid-angle pdfs of query results."""
    lib = load_library()
    rc = lib.dfg_directions(_stream(stream), wi.shape[0], _ptr(wi), _ptr(dir), _ptr(pdf))
    if rc != 0:
        raise DfgError(rc, "dfg_directions")


def scatter(wi, bsdf_pdf, bsdf_val, cos_theta, thp, bsdf_fraction=0.5, guided=None, mix=None, scatter_pdf=None,
            mis=None, alive=None, stream=None):
    """Mixture pdf and throughput update of the scatter step (thp updated in place)."""
    lib = load_library()
    rc = lib.dfg_scatter(_stream(stream), wi.shape[0], bsdf_fraction, _ptr(wi), _ptr(guided), _ptr(mix),
                         _ptr(bsdf_pdf), _ptr(bsdf_val), _ptr(cos_theta), _ptr(thp), _ptr(scatter_pdf), _ptr(mis),
                         _ptr(alive))
    if rc != 0:
        raise DfgError(rc, "dfg_scatter")


def sample_wi(depth, smooth, u_rr, u_mix, bsdf_wi, bsdf_pdf, bsdf_f, wi, rr_killed=None, guided=None, mix=None,
              thp=None, rr_prob=0.8, bsdf_fraction=0.3, guiding_enabled=True, max_guided_depth=6, stream=None):
    """The RR and guide-or-BSDF decision before the query (sampleWi)."""
    lib = load_library()
    rc = lib.dfg_sample_wi(_stream(stream), depth.shape[0], rr_prob, bsdf_fraction, int(guiding_enabled),
                           max_guided_depth, _ptr(depth), _ptr(smooth), _ptr(u_rr), _ptr(u_mix), _ptr(bsdf_wi),
                           _ptr(bsdf_pdf), _ptr(bsdf_f), _ptr(rr_killed), _ptr(guided), _ptr(mix), _ptr(wi), _ptr(thp))
    if rc != 0:
        raise DfgError(rc, "dfg_sample_wi")


def record_scatter(rec, sd, pixel, depth, alive=None, wi_dir=None, thp=None, scatter_pdf=None, mis=None,
                   bsdf_val=None, cos_theta=None, delta=None, rr_killed=None, train_enabled=True, max_train_depth=5,
                   pixel_offset=0, pixel_stride=1, mis_aware=True, product_with_cosine=False, stream=None):
    """Training records of a scattered (alive) or ended path vertex (handleIntersections)."""
    lib = load_library()
    r, h = records_struct(rec), shading_struct(sd)
    rc = lib.dfg_record_scatter(_stream(stream), pixel.shape[0], ctypes.byref(r), ctypes.byref(h), int(train_enabled),
                                max_train_depth, pixel_offset, pixel_stride, int(mis_aware), int(product_with_cosine),
                                _ptr(pixel), _ptr(depth), _ptr(rr_killed), _ptr(alive), _ptr(wi_dir), _ptr(thp),
                                _ptr(scatter_pdf), _ptr(mis), _ptr(bsdf_val), _ptr(cos_theta), _ptr(delta))
    if rc != 0:
        raise DfgError(rc, "dfg_record_scatter")


def record_radiance(rec, pixel, contrib, Le=None, miss=None, depth=None, train_enabled=True, max_train_depth=5,
                    pixel_offset=0, pixel_stride=1, stream=None):
    """Radiance reaching the recorded path vertices (handleEmissiveHit / handleMiss)."""
    lib = load_library()
    r = records_struct(rec)
    rc = lib.dfg_record_radiance(_stream(stream), pixel.shape[0], ctypes.byref(r), int(train_enabled),
                                 max_train_depth, pixel_offset, pixel_stride, _ptr(pixel), _ptr(depth), _ptr(contrib),
                                 _ptr(Le), _ptr(miss))
    if rc != 0:
        raise DfgError(rc, "dfg_record_radiance")


def prepare_features(pos, aabb, wo, normal, roughness, feat, stream=None):
    lib = load_library()
    code = lib.dfg_prepare_features(_stream(stream), pos.shape[0], _ptr(pos), _ptr(aabb), _ptr(wo),
                                    _ptr(normal), _ptr(roughness), _ptr(feat))
    if code != 0:
        raise DfgError(code, "dfg_prepare_features")
