"""Shared helpers for the GPU parity tests and the bench: synthetic inputs of
the shapes the guide consumes (seeded), and device-tensor plumbing."""
import math

import numpy as np


def df_features(n, seed, nrc=False):
    """DF feature rows [n][9] (or [n][15] with the NRC extras) as the reference's
    generate_inference_data_AR produces them (train.cu:334-362):
    pos U[0,1]^3; wo from dir2 ~ U[0,1]^2 read as (theta/pi, phi/2pi), warped
    (v+1)/2; normal2 ~ U[0,1]^2 (already theta/pi, phi/2pi); 1-exp(-r),
    r ~ U[0,1]; NRC extras diffuse3, specular3 ~ U[0,1]."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3), dtype=np.float32)
    d2 = rng.random((n, 2), dtype=np.float32)
    th, ph = d2[:, 0] * np.float32(math.pi), d2[:, 1] * np.float32(2 * math.pi)
    wo = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1).astype(np.float32)
    wo = (wo + 1.0) * 0.5
    nrm = rng.random((n, 2), dtype=np.float32)
    rough = 1.0 - np.exp(-rng.random((n, 1), dtype=np.float32))
    cols = [pos, wo, nrm, rough]
    if nrc:
        cols.append(rng.random((n, 6), dtype=np.float32))
    return np.ascontiguousarray(np.concatenate(cols, 1).astype(np.float32))


def plain_features(n, seed, dims=8):
    return np.random.default_rng(seed).random((n, dims), dtype=np.float32)


def uniforms(n, seed):
    return np.random.default_rng(seed).random((n, 4), dtype=np.float32)


def wi_init(n):
    w = np.zeros((n, 4), np.float32)
    w[:, 2:] = 1.0
    return w


def train_targets(n, seed):
    """GuidedOutput rows [B][9]: dir (u0,u1) of a direction uniform on the
    sphere via uniformSphereToSquare (math_utils.h:233-240), L ~ logNormal(0,2)
    per channel, wiPdf = 1/(4pi), bsdfVal = 1."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    phi = np.arctan2(v[:, 1], v[:, 0]) / (2 * math.pi)
    phi = np.where(phi < 0, phi + 1.0, phi)
    u0 = (1.0 - v[:, 2]) * 0.5
    L = np.exp(rng.normal(0.0, 2.0, (n, 3)))
    t = np.zeros((n, 9), np.float32)
    t[:, 0], t[:, 1] = u0, phi
    t[:, 2:5] = L
    t[:, 5] = 1.0 / (4 * math.pi)
    t[:, 6:9] = 1.0
    return t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def random_records(n_pixels, max_depth=5, seed=0, aabb=(-2.0, -1.0, 0.5, 3.0, 2.0, 4.0)):
    """Synthetic GuidedPixelStateBuffer contents (host numpy, SoA [depth][pixel]):
    positions inside the scene box, directions as spherical angles, radiance
    log-normal, with delta / miss vertices, zero pdfs, tiny throughputs and NaN
    radiance sprinkled in so that every skip rule of the reference fires."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array(aabb[:3], np.float32), np.array(aabb[3:], np.float32)
    D, N = max_depth, n_pixels
    sph = lambda: np.stack([rng.uniform(0, np.pi, (D, N)), rng.uniform(0, 2 * np.pi, (D, N))], -1).astype(np.float32)
    rec = {
        "cur_depth": rng.integers(0, D + 1, N).astype(np.uint32),
        "pos": (lo + (hi - lo) * rng.random((D, N, 3))).astype(np.float32),
        "dir": sph(), "wo": sph(), "normal": sph(),
        "roughness": rng.random((D, N)).astype(np.float32),
        "L": np.exp(rng.normal(0, 1, (D, N, 3))).astype(np.float32),
        "thp": rng.uniform(0.05, 1.0, (D, N, 3)).astype(np.float32),
        "bsdf_val": rng.uniform(0.0, 1.0, (D, N, 3)).astype(np.float32),
        "diffuse": rng.random((D, N, 3)).astype(np.float32),
        "specular": rng.random((D, N, 3)).astype(np.float32),
        "wi_pdf": rng.uniform(0.05, 2.0, (D, N)).astype(np.float32),
        "mis_weight": rng.uniform(0.0, 1.0, (D, N)).astype(np.float32),
        "delta": (rng.random((D, N)) < 0.1).astype(np.uint8),
        "miss": (rng.random((D, N)) < 0.1).astype(np.uint8),
    }
    rec["wi_pdf"][rng.random((D, N)) < 0.03] = 0.0
    rec["thp"][rng.random((D, N, 3)) < 0.03] = 1e-6
    rec["L"][rng.random((D, N, 3)) < 0.01] = np.nan
    return rec, np.array(aabb, np.float32)


def trained_scale_params(L, rng, gain=2.0, grid_amp=0.5):
    """fp16-representable parameters of one net at the scale training reaches
    (theta/phi logits of sigma ~1.5, grid entries far above the U[-1e-4, 1e-4]
    initialisation): Xavier-uniform matrices times `gain`, grid entries
    U[-grid_amp, grid_amp].  L is an oracle layout."""
    p = np.zeros(L["n_params"], np.float32)
    offs = list(L["layer_off"]) + [L["n_matrix"]]
    for i in range(len(offs) - 1):
        a, b = offs[i], offs[i + 1]
        fan_in = L["enc_width"] if i == 0 else 64
        fan_out = (b - a) // fan_in
        lim = gain * math.sqrt(6.0 / (fan_in + fan_out))
        p[a:b] = rng.uniform(-lim, lim, b - a)
    if L["grid"]:
        g0 = L["grid"]["param_begin"]
        p[g0:] = rng.uniform(-grid_amp, grid_amp, L["n_params"] - g0)
    return p.astype(np.float16).astype(np.float32)


def keep_grads(cfg):
    """A copy of the config whose context also stores the paired grids'
    gradient: their optimizer step otherwise consumes it inside the grid
    backward, and dfg_get_grad then refuses to return it."""
    import copy
    c = copy.deepcopy(cfg)
    c.setdefault("dfg", {})["keep_gradients"] = 1
    return c
