"""CPU tests that pin the oracle itself (closed forms, invariants, brute force,
finite differences).  The reference holds no golden vectors for this path, so
these are what make the oracle trustworthy before it checks the GPU."""
import math

import numpy as np
import pytest

import oracle as O
import dfguide as G
from gpu_util import df_features, uniforms, wi_init, trained_scale_params, train_targets


# ----------------------------------------------------------------------------
# deterministic exp / log / half rounding
# ----------------------------------------------------------------------------
def _ulps(a, exact):
    a = np.float32(a)
    return abs(float(a) - exact) / float(np.spacing(np.float32(exact)))


def test_det_exp_within_2ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87, 88, 20000), rng.uniform(-1, 1, 5000), [0.0, 1.0, -1.0, 88.0, -87.0]])
    worst = max(_ulps(O.det_expf(np.float32(x)), math.exp(float(np.float32(x)))) for x in xs)
    assert worst <= 2.0, worst
    assert O.det_expf(0.0) == 1.0
    assert O.det_expf(100.0) == np.inf and O.det_expf(-110.0) == 0.0
    # subnormal results stay within 2 ulp of the subnormal spacing
    for x in (-88.0, -95.5, -100.0, -103.0):
        ex = math.exp(x)
        assert abs(float(O.det_expf(x)) - ex) <= 2 * 2.0 ** -149


def test_det_log_within_2ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([10.0 ** rng.uniform(-37, 38, 20000), rng.uniform(0.5, 2.0, 5000), [1.0, 2.0, 64.0]])
    worst = 0.0
    for x in xs:
        x = np.float32(x)
        ex = math.log(float(x))
        if ex == 0.0:
            assert O.det_logf(x) == 0.0
            continue
        worst = max(worst, _ulps(O.det_logf(x), ex))
    assert worst <= 2.0, worst
    assert O.det_logf(0.0) == -np.inf and np.isnan(O.det_logf(-1.0))
    assert abs(float(O.det_logf(np.float32(1e-40))) - math.log(float(np.float32(1e-40)))) < 1e-5


def test_half_round_matches_ieee():
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-5, 2000), rng.normal(0, 3e4, 2000),
                         [65504.0, 65520.0, 70000.0, 2 ** -24, 2 ** -25, 3 * 2 ** -26, 0.0]]).astype(np.float32)
    for x in xs:
        assert np.float32(np.float16(x)) == O.half_round(x) or (np.isnan(O.half_round(x)))


# ----------------------------------------------------------------------------
# layouts: the numbers the reference's configurations produce
# ----------------------------------------------------------------------------
def test_reference_layout_numbers():
    th, ph, nrc = O.config_layouts(G.reference_config())
    assert (th["enc_width"], ph["enc_width"], nrc["enc_width"]) == (64, 80, 112)  # 60->64, 72->80, 102->112
    assert th["grid"]["offset"][-1] == 1033288                 # DenseGrid(8, F=4, base 8, s 1.405)
    assert th["grid"]["offset"][-1] * 4 == 4133152
    assert th["grid"]["res"][0] == 8 and th["grid"]["type"] == "dense"
    # the last nested encoding carries the padding to 16, SH pads in front (none here)
    parts = th["parts"]
    assert [p["width"] for p in parts] == [32, 16, 12] and parts[-1]["n_pad"] == 4
    assert th["n_matrix"] == 64 * 64 + 2 * 64 * 64 + 16 * 64
    assert ph["n_matrix"] == 64 * 80 + 2 * 64 * 64 + 32 * 64
    assert nrc["n_matrix"] == 64 * 112 + 4 * 64 * 64 + 16 * 64


def test_hash_layout_numbers():
    th, ph, _ = O.config_layouts(G.hash_config())
    g = th["grid"]
    assert g["offset"][-1] == 7114752 and g["offset"][-1] * 2 == 14229504
    assert g["res"] == [2 ** (l + 4) for l in range(16)]
    sizes = np.diff(g["offset"])
    assert list(sizes[:3]) == [16 ** 3, 32 ** 3, 64 ** 3] and all(s == 2 ** 19 for s in sizes[3:])
    assert th["enc_width"] == 64 and ph["enc_width"] == 80


def test_plain_head_layout():
    th, ph, _ = O.config_layouts(G.plain_head_config())
    assert th["enc_width"] == 16 and ph["enc_width"] == 32   # 8 -> 16; 8 + 16 -> 32
    assert th["n_matrix"] == 64 * 16 + 2 * 4096 + 32 * 64


# ----------------------------------------------------------------------------
# grid indexing: brute force and the uint32 stride wrap of the reference
# ----------------------------------------------------------------------------
def test_grid_index_dense_and_hash_brute_force():
    lib = O.lib()
    res, T = 9, 1024  # 729 <= 1024: dense
    for x, y, z in np.random.default_rng(3).integers(0, 12, (500, 3)):
        assert lib.orc_grid_index3(1, T, res, x, y, z) == (x + y * res + z * res * res) % T
    res = 20  # 8000 > 1024: hashed
    for x, y, z in np.random.default_rng(4).integers(0, 40, (500, 3)):
        h = (int(x) ^ ((int(y) * 2654435761) & 0xffffffff) ^ ((int(z) * 805459861) & 0xffffffff)) % T
        assert lib.orc_grid_index3(1, T, res, x, y, z) == h
        assert lib.orc_grid_index3(0, T, res, x, y, z) == (x + y * res + z * res * res) % T  # dense never hashes


def test_grid_index_uint32_wrap_levels():
    """With res = 2^16..2^19 and T = 2^19, res^2 wraps to 0 in the reference's
    uint32 stride, so those levels keep the dense formula: index = (x + y res) mod T,
    independent of z."""
    lib = O.lib()
    T = 2 ** 19
    rng = np.random.default_rng(5)
    for res in (2 ** 16, 2 ** 17, 2 ** 18, 2 ** 19):
        for x, y, z in rng.integers(0, 2 ** 19, (200, 3)):
            expect = ((int(x) + int(y) * res) & 0xffffffff) % T
            assert lib.orc_grid_index3(1, T, res, x, y, z) == expect
    # res = 2^15: res^2 = 2^30 > T, so the hash is used
    res = 2 ** 15
    x, y, z = 123, 456, 789
    assert lib.orc_grid_index3(1, T, res, x, y, z) == (x ^ ((y * 2654435761) & 0xffffffff) ^ ((z * 805459861) & 0xffffffff)) % T


def _near_vertex_positions(scales, n_per_level, rng):
    """float32 inputs x within a few ulps of (k - 0.5) / scale for random
    integers k, i.e. positions x * scale + 0.5 within an ulp or two of a cell
    boundary, on every level's scale."""
    out = []
    for s in scales:
        k = rng.integers(1, int(s), n_per_level)
        x = ((k - 0.5) / s).astype(np.float32)
        for step in range(-3, 4):
            y = x.copy()
            for _ in range(abs(step)):
                y = np.nextafter(y, np.float32(np.inf if step > 0 else -np.inf))
            out.append(y)
    return np.concatenate(out)


def test_grid_positions_are_exact():
    """The cell of every level is floor(x * scale + 0.5) of the exact real
    product, also within an ulp of a boundary, where the reference's fp32
    rounding of the position lands in the neighbouring cell: the FP32 mode's
    corner indices equal those composed in float64 (exact for float inputs)
    through the brute-force-pinned index function."""
    cfg = G.hash_config(with_nrc=False)
    th, _, _ = O.config_layouts(cfg)
    g = th["grid"]
    rng = np.random.default_rng(11)
    scales = [float(s) for s in g["scale"]]
    xs = _near_vertex_positions(scales, 40, rng)
    pos = np.stack([xs, rng.permutation(xs), rng.permutation(xs)], 1).astype(np.float32)
    feat = np.zeros((pos.shape[0], 9), np.float32)
    feat[:, :3] = pos
    got = O.grid_index(th, feat)
    lib = O.lib()
    rounded_differs = 0
    for l in range(g["levels"]):
        s = np.float64(np.float32(scales[l]))
        exact = np.floor(pos.astype(np.float64) * s + 0.5).astype(np.int64)
        rounded = np.floor(pos * np.float32(s) + np.float32(0.5)).astype(np.int64)
        rounded_differs += int(np.any(exact != rounded, axis=1).sum())
        hsize = g["offset"][l + 1] - g["offset"][l]
        for i in range(0, pos.shape[0], 7):
            for k in range(8):
                c = [int(exact[i, d] + ((k >> d) & 1)) & 0xffffffff for d in range(3)]
                assert got[i, l, k] == lib.orc_grid_index3(1, hsize, g["res"][l], *c), (i, l, k)
    # the probe set really contains positions that fp32 rounding puts in another cell
    assert rounded_differs > 100


def test_fp32_query_tracks_fp64_on_trained_scale_weights():
    """With exact grid positions the FP32 restatement stays within 1e-5
    (median) of FP64 on weights of trained scale (logits of sigma ~1.5,
    grid entries U[-0.5, 0.5]), where fp32-rounded positions give ~1e-3."""
    cfg = G.hash_config(with_nrc=False)
    th, ph, _ = O.config_layouts(cfg)
    rng = np.random.default_rng(5)
    pt, pp = trained_scale_params(th, rng), trained_scale_params(ph, rng)
    n = 2048
    feat, u = df_features(n, 1), uniforms(n, 2)
    o32, _ = O.query_f32_omp(th, ph, pt, pp, feat, 9, u, wi_init(n), 1)
    at = wi_init(n)
    at[:, :2] = o32[:, :2]
    o64 = O.query_f64(th, ph, pt, pp, feat, 9, u, at, 1, guided=np.zeros(n, np.uint8))
    ref = o64[:, 2].astype(np.float64) * o64[:, 3]
    rel = np.abs(o32[:, 2].astype(np.float64) * o32[:, 3] - ref) / ref
    assert np.median(rel) <= 1e-5, np.median(rel)
    assert np.percentile(rel, 99) <= 1e-3, np.percentile(rel, 99)


def _grid_cfg(interp="Linear", gtype="HashGrid", levels=4, F=2, log2T=10):
    return {"otype": "Composite", "nested": [
        {"otype": gtype, "n_levels": levels, "n_features_per_level": F, "log2_hashmap_size": log2T,
         "base_resolution": 4, "per_level_scale": 1.5, "interpolation": interp, "n_dims_to_encode": 3}]}


@pytest.mark.parametrize("interp", ["Linear", "Smoothstep", "Nearest"])
def test_grid_interpolation_properties(interp):
    L = O.layout(_grid_cfg(interp), {"n_hidden_layers": 1}, 3, 16)
    rng = np.random.default_rng(6)
    params = rng.normal(0, 1, L["n_params"])
    g = L["grid"]
    # constant table -> constant output (weights sum to one)
    pc = params.copy()
    pc[g["param_begin"]:] = 0.75
    x = rng.random((300, 3))
    e = O.encode_f64(L, pc, x)
    assert np.allclose(e[:, :g["levels"] * g["F"]], 0.75, atol=1e-12)
    # linear in the table entries
    a, b = rng.normal(0, 1, L["n_params"]), rng.normal(0, 1, L["n_params"])
    ea, eb, eab = O.encode_f64(L, a, x), O.encode_f64(L, b, x), O.encode_f64(L, 2 * a - 3 * b, x)
    assert np.allclose(eab[:, :8], 2 * ea[:, :8] - 3 * eb[:, :8], atol=1e-10)
    # at a vertex (frac == 0) the value is the table entry of that vertex
    # (float inputs whose exact position x * scale + 0.5 lies at the vertex or
    # at most a few float ulps above it, so that the vertex is corner 0)
    lvl0_scale = float(g["scale"][0])
    cell = np.array([1, 2, 0])
    xv = ((cell - 0.5) / lvl0_scale).astype(np.float32)
    for d in range(3):
        while float(xv[d]) * lvl0_scale + 0.5 < cell[d]:
            xv[d] = np.nextafter(xv[d], np.float32(np.inf))
    frac = xv.astype(np.float64) * lvl0_scale + 0.5 - cell
    assert np.all((frac >= 0) & (frac < 1e-5))
    xv = xv[None, :]
    idx = O.grid_index(L, xv)[0, 0, 0]
    ev = O.encode_f64(L, params, xv.astype(np.float64))
    t0 = params[g["param_begin"]:]
    assert np.allclose(ev[0, :2], t0[idx * 2: idx * 2 + 2], atol=1e-4)


def test_smoothstep_closed_form():
    L = O.layout(_grid_cfg("Smoothstep", levels=1), {"n_hidden_layers": 1}, 3, 16)
    g = L["grid"]
    params = np.zeros(L["n_params"])
    # table holding the vertex x coordinate in feature 0 -> output is the smoothstep of the x fraction
    res = g["res"][0]
    for xi in range(res + 1):
        for yi in range(res + 1):
            for zi in range(res + 1):
                idx = O.lib().orc_grid_index3(1, g["offset"][1], res, xi, yi, zi)
                params[g["param_begin"] + idx * 2] = xi
    s = float(g["scale"][0])
    for t in (0.1, 0.3, 0.5, 0.9):
        x = np.array([[(1 + t - 0.5) / s, 0.3, 0.6]])
        e = O.encode_f64(L, params, x)[0, 0]
        frac = x[0, 0] * s + 0.5 - 1
        assert abs(e - (1 + frac * frac * (3 - 2 * frac))) < 1e-9


# ----------------------------------------------------------------------------
# SH / OneBlob / TriangleWave / Frequency / Identity
# ----------------------------------------------------------------------------
def _single(enc, in_dims):
    return O.layout(enc, {"n_hidden_layers": 1}, in_dims, 16)


def test_sh_orthonormal_by_quadrature():
    L = _single({"otype": "SphericalHarmonics", "degree": 4}, 3)
    ct, wt = np.polynomial.legendre.leggauss(12)
    nphi = 24
    phis = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    pts, w = [], []
    for c, wc in zip(ct, wt):
        s = math.sqrt(1 - c * c)
        for p in phis:
            pts.append([s * math.cos(p), s * math.sin(p), c])
            w.append(wc * 2 * math.pi / nphi)
    pts = (np.array(pts) + 1) / 2
    Y = O.encode_f64(L, np.zeros(L["n_params"]), pts)[:, :16]
    gram = (Y * np.array(w)[:, None]).T @ Y
    assert np.allclose(gram, np.eye(16), atol=1e-10)


def test_oneblob_partition_of_unity():
    L = _single({"otype": "OneBlob", "n_bins": 4}, 1)
    x = np.random.default_rng(7).random((500, 1))
    e = O.encode_f64(L, np.zeros(L["n_params"]), x)[:, :4]
    assert np.allclose(e.sum(1), 1.0, atol=1e-12)
    assert np.all(e >= -1e-12)
    # symmetric point: x = 0.5 sits on the boundary of bins 1 and 2
    e = O.encode_f64(L, np.zeros(L["n_params"]), np.array([[0.5]]))[0, :4]
    assert abs(e[1] - e[2]) < 1e-12 and abs(e[0] - e[3]) < 1e-12


def test_triangle_wave_and_frequency_closed_forms():
    L = _single({"otype": "TriangleWave", "n_frequencies": 12}, 1)
    x = np.random.default_rng(8).random((50, 1))
    e = O.encode_f64(L, np.zeros(L["n_params"]), x)[:, :12]
    for f in range(12):
        v = x[:, 0] * 2.0 ** (f - 1) + f * 0.25
        assert np.allclose(e[:, f], np.abs(v - np.floor(v) - 0.5) * 4 - 1, atol=1e-12)
    L = _single({"otype": "Frequency", "n_frequencies": 3}, 2)
    x = np.random.default_rng(9).random((50, 2))
    e = O.encode_f64(L, np.zeros(L["n_params"]), x)[:, :12]
    for d in range(2):
        for j in range(6):
            f = j // 2
            ref = np.sin(2.0 ** f * math.pi * x[:, d]) if j % 2 == 0 else np.cos(2.0 ** f * math.pi * x[:, d])
            assert np.allclose(e[:, d * 6 + j], ref, atol=1e-12)
    L = _single({"otype": "Identity", "scale": 2.0, "offset": -1.0}, 3)
    e = O.encode_f64(L, np.zeros(L["n_params"]), x[:, :1].repeat(3, 1))
    assert np.allclose(e[:, :3], 2 * x[:, :1] - 1) and np.all(e[:, 3:] == 1.0)   # pads are 1.0


# ----------------------------------------------------------------------------
# DF head: softmax, sampler, pdf
# ----------------------------------------------------------------------------
def _cdf_at(bins, x):
    """Closed-form CDF of the piecewise-linear density with constant end half-bins."""
    N = bins.size
    grid = np.concatenate([[0.0], (np.arange(N) + 0.5) / N, [1.0]])
    vals = np.concatenate([[bins[0]], bins, [bins[-1]]])
    c = 0.0
    for k in range(len(grid) - 1):
        a, b = grid[k], grid[k + 1]
        if x <= a:
            break
        t = min(x, b)
        slope = (vals[k + 1] - vals[k]) / (b - a)
        c += vals[k] * (t - a) + 0.5 * slope * (t - a) ** 2
    return c


@pytest.mark.parametrize("N", [2, 4, 16, 32, 64])
def test_sampler_inverts_the_cdf(N):
    rng = np.random.default_rng(10 + N)
    n = 400
    z = rng.normal(0, 1.5, (n, N))
    u = rng.random(n)
    x, pdf, bins, seg = O.head(z, 1, u, rng.random(n))
    assert np.allclose(bins.sum(1) / N, 1.0, atol=1e-12)
    for i in range(n):
        if abs(bins[i, min(seg[i], N - 1)] - bins[i, max(seg[i] - 1, 0)]) * N * 0.5 < 1e-5:
            continue  # linear-fallback segment, exact only to the A ~ 0 approximation
        assert abs(_cdf_at(bins[i], x[i]) - u[i]) < 1e-9, (i, x[i], u[i])


def test_uniform_logits_sample_is_identity():
    n = 1000
    u = np.random.default_rng(11).random(n) * (1 - 1e-5)
    for N in (16, 32, 64):
        x, pdf, bins, _ = O.head(np.zeros((n, N)), 1, u, np.zeros(n))
        assert np.allclose(bins, 1.0) and np.allclose(x, u, atol=1e-12) and np.allclose(pdf, 1.0)


def test_pdf_integrates_to_one_and_end_bins_constant():
    rng = np.random.default_rng(12)
    for N in (16, 32):
        z = rng.normal(0, 2, (1, N))
        _, _, bins, _ = O.head(z, 1, np.array([0.5]), np.array([0.5]))
        xs = (np.arange(200000) + 0.5) / 200000
        p = np.array([O.lib().orc_pdf_linear_f64(bins[0].ctypes.data_as(O.ctypes.c_void_p), N, float(x))
                      for x in xs[::50]])
        assert abs(p.mean() - 1.0) < 2e-3
        f = lambda x: O.lib().orc_pdf_linear_f64(bins[0].ctypes.data_as(O.ctypes.c_void_p), N, x)
        assert f(0.0) == f(0.25 / N) == bins[0, 0]
        assert f(1.0 - 1e-9) == f(1.0 - 0.25 / N) == bins[0, -1]
        assert abs(f((3 + 0.5) / N) - bins[0, 3]) < 1e-12     # value at a bin centre


def test_pdf_is_total_beyond_the_unit_interval():
    """The pdf of a given square coordinate is defined for every float: the
    constant end half-bins extend beyond [0, 1] (the reference's own indices
    read out of bounds below the first bin centre's half-bin), infinities
    take the end bins and NaN gives NaN; the nearest mode clamps too."""
    rng = np.random.default_rng(31)
    for N in (16, 32, 64):
        _, _, bins, _ = O.head(rng.normal(0, 2, (1, N)), 1, np.array([0.5]), np.array([0.5]))
        f = lambda x: O.lib().orc_pdf_linear_f64(bins[0].ctypes.data_as(O.ctypes.c_void_p), N, x)
        for x in (-1e-9, -0.3 / N, -1.0 / N, -0.7, -5.0, -1e30, -np.inf):
            assert f(x) == bins[0, 0], x
        for x in (1.0, 1.0 + 0.3 / N, 1.7, 1e30, np.inf):
            assert f(x) == bins[0, -1], x
        assert np.isnan(f(np.nan))
        # inside the interval: unchanged (interior linear interpolation)
        x = (5 + 0.5 + 0.25) / N   # a quarter of the way from bin 5's centre to bin 6's
        assert abs(f(x) - (0.75 * bins[0, 5] + 0.25 * bins[0, 6])) < 1e-12
    # the head with given coordinates (unguided rows), FP64 and FP32 modes
    N = 32
    z = rng.normal(0, 1.5, (8, N))
    xg = np.array([-2.0, -1e-6, 0.0, 0.5, 1.0, 3.0, np.inf, -np.inf])
    for fp32 in (False, True):
        for linear in (0, 1):
            x, pdf, bins, _ = O.head(z, linear, np.zeros(8), np.zeros(8), sample=np.zeros(8, np.uint8),
                                     x_given=xg, fp32=fp32)
            assert np.all(np.isfinite(pdf)) and np.all(pdf > 0)
            assert pdf[0] == bins[0, 0] and pdf[1] == bins[1, 0] and pdf[6] == bins[6, -1] and pdf[7] == bins[7, 0]


def test_samples_follow_the_pdf_chi2():
    rng = np.random.default_rng(13)
    N = 16
    z = np.tile(rng.normal(0, 1.2, N), (200000, 1))
    u = rng.random(200000)
    x, _, bins, _ = O.head(z, 1, u, rng.random(200000))
    hist, edges = np.histogram(x, bins=64, range=(0, 1))
    expect = np.array([_cdf_at(bins[0], b) - _cdf_at(bins[0], a) for a, b in zip(edges[:-1], edges[1:])]) * x.size
    chi2 = ((hist - expect) ** 2 / expect).sum()
    assert chi2 < 120, chi2   # 63 dof; p ~ 1e-5 threshold


def test_nearest_mode():
    N = 8
    z = np.random.default_rng(14).normal(0, 1, (300, N))
    u = np.random.default_rng(15).random(300)
    x, pdf, bins, seg = O.head(z, 0, u, np.zeros(300))
    c = np.cumsum(bins / N, 1)
    for i in range(300):
        k = int(np.searchsorted(c[i], u[i], side="right"))
        assert seg[i] == k
        lo = c[i, k - 1] if k else 0.0
        assert abs(x[i] - (k + (u[i] - lo) / (bins[i, k] / N)) / N) < 1e-12
        assert pdf[i] == bins[i, min(int(x[i] * N), N - 1)]


def test_fp32_head_tracks_fp64():
    rng = np.random.default_rng(16)
    n, N = 5000, 32
    z = rng.normal(0, 1, (n, N))
    u, ue = rng.random(n), rng.random(n)
    x64, p64, _, s64 = O.head(z, 1, u, ue)
    x32, p32, _, s32 = O.head(z.astype(np.float32), 1, u.astype(np.float32), ue.astype(np.float32), fp32=True)
    assert np.mean(s64 == s32) > 0.999
    assert np.median(np.abs(x32 - x64)) < 1e-6 and np.median(np.abs(p32 / p64 - 1)) < 1e-5


def _global_form_root(bins, N, seg, u):
    """The reference's quadratic in global coordinates (train.h:453-506) in
    float64: A x^2 + B x + C = 0 with the root selected by the sign of A; NaN
    where the reference takes its linear or extra-uniform fallback."""
    i = seg
    vim1 = i - 1
    v = vim1 + 0.5
    if vim1 < 0 or i > N - 1:
        return np.nan
    m = (bins[i] - bins[vim1]) * N
    B = bins[vim1] * (1 + v) - bins[i] * v
    A = m * 0.5
    if abs(A) < 1e-5:
        return np.nan
    lo = bins[0] * 0.5 / N + sum((bins[k - 1] + bins[k]) / (2 * N) for k in range(1, i))
    C = -A * (v / N) ** 2 - B * v / N + lo - u
    delta = B * B - 4 * A * C
    if delta < 1e-5:
        return np.nan
    s0, s1 = (-B + math.sqrt(delta)) / (2 * A), (-B - math.sqrt(delta)) / (2 * A)
    return max(s0, s1) if A >= 0 else min(s0, s1)


@pytest.mark.parametrize("N", [16, 32, 64])
def test_sampler_root_equals_the_reference_global_form(N):
    """The sampler takes the root in the segment's own coordinate; in float64
    that is the reference's global-coordinate root to rounding, on peaked
    bins where both roots of the quadratic lie near the segment."""
    rng = np.random.default_rng(30 + N)
    n = 3000
    z = rng.normal(0, 2.0, (n, N))
    u = rng.random(n)
    x, _, bins, seg = O.head(z, 1, u, rng.random(n))
    checked = 0
    for i in range(n):
        r = _global_form_root(bins[i], N, seg[i], u[i])
        if np.isnan(r):
            continue
        checked += 1
        assert abs(x[i] - min(r, 1 - 1e-5)) < 1e-10, (i, x[i], r)
    assert checked > n // 2


def test_fp32_sample_tracks_fp64_on_peaked_bins():
    """With the local-coordinate root the FP32 sample stays within an ulp or
    so of FP64 (median) on the logit spread of trained nets (sigma 2, 64
    bins); the global-coordinate form loses 25x more."""
    rng = np.random.default_rng(17)
    n, N = 5000, 64
    z = rng.normal(0, 2.0, (n, N))
    u, ue = rng.random(n), rng.random(n)
    x64, _, _, s64 = O.head(z, 1, u, ue)
    x32, _, _, s32 = O.head(z.astype(np.float32), 1, u.astype(np.float32), ue.astype(np.float32), fp32=True)
    same = s64 == s32
    assert np.mean(same) > 0.999
    dx = np.abs(x32 - x64)[same]
    # global form: median 3.6e-7, p99 7.7e-6; local form: 1.5e-8, 8.7e-7
    assert np.median(dx) < 5e-8 and np.percentile(dx, 99) < 3e-6, (np.median(dx), np.percentile(dx, 99))


def test_theta_input():
    lib = O.lib()
    for th in np.random.default_rng(17).random(1000):
        assert abs(lib.orc_theta_input_f64(th, 16, 1) - th) < 1e-12
        assert lib.orc_theta_input_f64(th, 16, 0) == math.floor(th * 16) / 16


# ----------------------------------------------------------------------------
# training: finite differences of the FP64 loss vs the oracle's analytic gradient
# ----------------------------------------------------------------------------
def _small_train_case(cfg, D, B, seed):
    from gpu_util import df_features, plain_features, train_targets
    th, ph, _ = O.config_layouts(cfg)
    rng = np.random.default_rng(seed)
    pt = rng.normal(0, 0.3, th["n_params"])
    pp = rng.normal(0, 0.3, ph["n_params"])
    if th["grid"]:
        pt[th["n_matrix"]:] = rng.normal(0, 0.5, th["n_params"] - th["n_matrix"])
        pp[ph["n_matrix"]:] = rng.normal(0, 0.5, ph["n_params"] - ph["n_matrix"])
    feat = plain_features(B, seed) if D == 8 else df_features(B, seed)
    tgt = train_targets(B, seed + 1)
    return th, ph, pt, pp, feat, tgt


@pytest.mark.parametrize("cfgname", ["plain", "grid"])
def test_df_gradient_finite_differences(cfgname):
    if cfgname == "plain":
        cfg, D = G.plain_head_config(), 8
    else:
        cfg, D = G.hash_config(with_nrc=False, n_levels=4, log2_hashmap_size=8), 9
    th, ph, pt, pp, feat, tgt = _small_train_case(cfg, D, 16, 20)
    loss, gt, gp = O.df_train_f64(th, ph, pt, pp, feat, D, tgt, 1)
    assert np.isfinite(loss)
    rng = np.random.default_rng(21)
    h = 1e-6
    for which, params, grad, L in ((0, pt, gt, th), (1, pp, gp, ph)):
        cand = list(rng.choice(L["n_matrix"], 12, replace=False))
        if L["grid"]:
            nz = np.nonzero(grad[L["n_matrix"]:])[0] + L["n_matrix"]
            cand += list(rng.choice(nz, 8, replace=False))
        for i in cand:
            p1, p2 = params.copy(), params.copy()
            p1[i] += h
            p2[i] -= h
            a = (pt, pp)
            l1 = O.df_train_f64(th, ph, *( (p1, pp) if which == 0 else (pt, p1) ), feat, D, tgt, 1, grads=False)[0]
            l2 = O.df_train_f64(th, ph, *( (p2, pp) if which == 0 else (pt, p2) ), feat, D, tgt, 1, grads=False)[0]
            fd = (l1 - l2) / (2 * h)
            assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(grad[i])) + 1e-7, (which, i, fd, grad[i])


def test_nrc_gradient_finite_differences():
    from gpu_util import df_features
    cfg = G.hash_config(with_nrc=True)
    _, _, L = O.config_layouts(cfg)
    rng = np.random.default_rng(22)
    P = rng.normal(0, 0.2, L["n_params"])
    feat = df_features(12, 23, nrc=True)
    tgt = rng.random((12, 3)).astype(np.float32)
    loss, g = O.nrc_train_f64(L, P, feat, tgt)
    # the reference treats the luminance as a constant; finite differences of the
    # loss with the denominator frozen at the current prediction
    def frozen_loss(p):
        o = O.mlp_f64(L, p, O.encode_f64(L, p, feat[:, :15].astype(np.float64)))
        o0 = O.mlp_f64(L, P, O.encode_f64(L, P, feat[:, :15].astype(np.float64)))
        lum = 0.299 * o0[:, 0] + 0.587 * o0[:, 1] + 0.114 * o0[:, 2]
        return (((o[:, :3] - tgt) ** 2).sum(1) / (lum ** 2 + 0.01)).sum() / (3 * 12)
    # the oracle reports sum(values) / B (nrc.cpp:166); its gradient is d sum(values)
    B = 12
    assert abs(frozen_loss(P) - loss * B) <= 1e-12 * max(1.0, abs(loss * B))
    for i in rng.choice(L["n_matrix"], 10, replace=False):
        p1, p2 = P.copy(), P.copy()
        p1[i] += 1e-6
        p2[i] -= 1e-6
        fd = (frozen_loss(p1) - frozen_loss(p2)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


# ----------------------------------------------------------------------------
# Adam + EMA closed forms
# ----------------------------------------------------------------------------
def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 0.0, 0.0, 1e-3])
    w = np.array([1.0, -1.0, 0.5, 0.25, 2.0])
    n_matrix = 3  # params 3, 4 are non-matrix (grid)
    lr, b1, b2, eps, l2 = 1e-2, 0.9, 0.99, 1e-15, 1e-6
    master, m, v, steps, _ = O.adam_ema_f64(n_matrix, lr, b1, b2, eps, l2, 0.95, 1, g, w, np.zeros(5), np.zeros(5),
                                            np.zeros(5, np.uint32))
    ge = g + np.where(np.arange(5) < n_matrix, l2 * w, 0.0)
    # first step: m = (1-b1) g, v = (1-b2) g^2, lr_t = lr sqrt(1-b2)/(1-b1) -> update = lr * sign(g)
    for i in (0, 1, 2, 4):
        step = lr * math.sqrt(1 - b2) / (1 - b1) * ((1 - b1) * ge[i]) / (math.sqrt((1 - b2) * ge[i] ** 2) + eps)
        assert abs(master[i] - (w[i] - step)) < 1e-12
    assert master[3] == w[3] and steps[3] == 0 and steps[2] == 1  # zero-grad non-matrix skipped; matrix not
    assert list(steps) == [1, 1, 1, 0, 1]


def test_ema_debias_identity():
    n = 7
    w = np.random.default_rng(24).normal(0, 1, n)
    ema = np.zeros(n)
    for step in range(1, 6):
        _, _, _, _, ema = O.adam_ema_f64(0, 0.0, 0.9, 0.99, 1e-8, 0.0, 0.95, step, np.zeros(n), w, np.zeros(n),
                                         np.zeros(n), np.zeros(n, np.uint32), w_lowp=w, ema=ema)
        assert np.allclose(ema, w, atol=1e-12)   # constant weights stay constant after debiasing


# ----------------------------------------------------------------------------
# MLP and whole-query composition against plain numpy
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("cfgname", ["hash", "reference"])
def test_mlp_is_a_relu_chain_of_row_major_matrices(cfgname):
    """The oracle's MLP equals numpy matmuls over the reference parameter
    layout (first, hidden..., last layer; each row-major [out][in]; ReLU on
    hidden layers, linear output) — a transposed or misplaced matrix fails."""
    cfg = G.hash_config(log2_hashmap_size=12) if cfgname == "hash" else G.reference_config()
    for L in O.config_layouts(cfg):
        if L is None:
            continue
        rng = np.random.default_rng(41)
        p = rng.normal(0, 0.2, L["n_params"])
        x = rng.normal(0, 1, (64, L["enc_width"]))
        h = x
        off = L["layer_off"]
        for l in range(L["n_hidden"] + 1):
            rows = L["n_out_pad"] if l == L["n_hidden"] else 64
            W = p[off[l]:off[l + 1]].reshape(rows, h.shape[1])
            h = h @ W.T
            if l < L["n_hidden"]:
                h = np.maximum(h, 0.0)
        assert np.allclose(O.mlp_f64(L, p, x), h, rtol=1e-12, atol=1e-12)


def test_query_composes_encode_mlp_head():
    """query_f64 = theta head of mlp(encode(feat)), then the phi head of
    mlp(encode(feat, theta_input(x_theta))), pdfs multiplied into wi."""
    from gpu_util import df_features
    cfg = G.hash_config(log2_hashmap_size=12, with_nrc=False)
    th, ph, _ = O.config_layouts(cfg)
    rng = np.random.default_rng(43)
    # query_f64 takes fp32 parameters (as the library holds them)
    pt = rng.normal(0, 0.3, th["n_params"]).astype(np.float32).astype(np.float64)
    pp = rng.normal(0, 0.3, ph["n_params"]).astype(np.float32).astype(np.float64)
    n = 256
    feat = df_features(n, 44)
    u = rng.random((n, 4))
    wi = np.concatenate([np.zeros((n, 2)), np.ones((n, 2))], 1).astype(np.float32)
    q = O.query_f64(th, ph, pt, pp, feat, 9, u.astype(np.float32), wi, 1)
    f64 = feat.astype(np.float64)
    u32 = u.astype(np.float32).astype(np.float64)
    lt = O.mlp_f64(th, pt, O.encode_f64(th, pt, f64))[:, :32]
    x0, p0, _, _ = O.head(lt, 1, u32[:, 0], u32[:, 2])
    t_in = np.array([O.lib().orc_theta_input_f64(float(v), 32, 1) for v in x0])
    lp = O.mlp_f64(ph, pp, O.encode_f64(ph, pp, np.concatenate([f64, t_in[:, None]], 1)))[:, :32]
    x1, p1, _, _ = O.head(lp, 1, u32[:, 1], u32[:, 3])
    assert np.allclose(q[:, 0], x0, rtol=0, atol=1e-6) and np.allclose(q[:, 1], x1, rtol=0, atol=1e-6)
    assert np.allclose(q[:, 2], p0, rtol=1e-6) and np.allclose(q[:, 3], p1, rtol=1e-6)


# ----------------------------------------------------------------------------
# training records -> rows, batch schedules
# ----------------------------------------------------------------------------
def test_sphere_square_maps_round_trip():
    rng = np.random.default_rng(50)
    u = rng.random((1000, 2)) * 0.998 + 0.001
    v = O.square_to_uniform_sphere(u)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    assert np.allclose(O.uniform_sphere_to_square(v), u, atol=1e-12)
    assert np.allclose(O.square_to_uniform_sphere(np.array([[0.0, 0.3], [1.0, 0.7], [0.5, 0.0], [0.5, 0.25]])),
                       [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0]], atol=1e-12)
    # sphericalToCartesian: theta from +z, phi from +x
    assert np.allclose(O.spherical_to_cartesian(np.array([0.0, np.pi / 2, np.pi / 2]), np.array([1.0, 0.0, np.pi / 2])),
                       [[0, 0, 1], [1, 0, 0], [0, 1, 0]], atol=1e-12)


def test_spatial_normalisation_closed_form():
    aabb = np.array([-2.0, -1.0, 0.5, 3.0, 2.0, 4.0])
    lo, hi = aabb[:3], aabb[3:]
    assert np.allclose(O.normalize_spatial((lo + hi) / 2, aabb), 0.5)
    d = np.linalg.norm(hi - lo) * 0.005
    assert np.allclose(O.normalize_spatial(lo, aabb), 0.5 - (hi - lo) / 2 / (hi - lo + 2 * d))
    assert np.allclose(O.normalize_spatial(hi, aabb), 0.5 + (hi - lo) / 2 / (hi - lo + 2 * d))


def test_training_rows_follow_the_skip_and_radiance_rules():
    from gpu_util import random_records
    rec, aabb = random_records(40, seed=51)
    f, t = O.training_targets_f64(rec, 40, 0, 1, aabb)
    # count by the rules, written out directly
    expect, rows = 0, []
    for p in range(40):
        for k in range(min(int(rec["cur_depth"][p]), 5)):
            if rec["delta"][k, p] or rec["wi_pdf"][k, p] == 0:
                continue
            thp, L = rec["thp"][k, p].astype(float), rec["L"][k, p].astype(float)
            loc = np.where(thp > 1e-5, L / thp, 0.0) * float(rec["mis_weight"][k, p])
            if np.isnan(loc).any():
                continue
            expect += 1
            rows.append(loc * rec["bsdf_val"][k, p])
    assert f.shape == (expect, 9) and t.shape == (expect, 9)
    assert np.allclose(t[:, 2:5], np.array(rows)) and np.all(t[:, 6:] == 0)
    assert np.all((t[:, :2] >= 0) & (t[:, :2] <= 1)) and np.all((f[:, :3] > 0) & (f[:, :3] < 1))
    fn, tn = O.nrc_training_targets_f64(rec, 40, 0, 1, aabb)
    assert fn.shape[1] == 15 and tn.shape[1] == 3 and len(fn) < len(f) + 40


def test_batch_schedules():
    sched, planned = O.df_batches(200000, 524288, 5)
    assert planned == 1 and sched == [(0, 199936)]
    sched, planned = O.df_batches(600000, 524288, 5)
    assert planned == 2 and sched == [(0, 524288), (524288, 75648)]
    sched, planned = O.df_batches(530000, 524288, 5)
    assert planned == 2 and sched == [(0, 524288)]            # the 5,712-row tail is under 65,536
    sched, planned = O.df_batches(50000, 524288, 5)
    assert planned == 1 and sched == []
    sched, planned = O.nrc_batches(1000000, 1048576, 4)
    assert planned == 4 and [b for _, b in sched] == [249984] * 4
    assert [o for o, _ in sched] == [0, 249984, 499968, 749952]


def test_scatter_mixture_closed_forms():
    rng = np.random.default_rng(60)
    n = 64
    wi = np.concatenate([rng.random((n, 2)), rng.uniform(0.1, 3, (n, 2))], 1)
    pb, bv, c, thp = rng.uniform(0.1, 2, n), rng.random((n, 3)), rng.random(n), rng.random((n, 3))
    pg = wi[:, 2] * wi[:, 3] * np.float64(np.float32(1 / (4 * np.pi)))
    ones = np.ones(n, np.uint8)
    # f = 0: guided rows scatter with the guide alone and carry MIS weight 1
    t, p, mis, alive = O.scatter_f64(0.0, wi, ones, None, pb, bv, c, thp)
    assert np.allclose(p, pg) and np.allclose(mis, 1.0) and alive.all()
    assert np.allclose(t, thp * bv * c[:, None] / (pg[:, None] + 1e-5))
    # f = 1, unguided, no mixing: the BSDF alone, MIS weight 0
    t, p, mis, _ = O.scatter_f64(1.0, wi, None, None, pb, bv, c, thp)
    assert np.allclose(p, pb) and np.allclose(mis, 0.0)
    # a mixed unguided row equals a guided row (same mixture)
    _, p1, _, _ = O.scatter_f64(0.3, wi, None, ones, pb, bv, c, thp)
    _, p2, _, _ = O.scatter_f64(0.3, wi, ones, None, pb, bv, c, thp)
    assert np.allclose(p1, p2)
    # a zero BSDF value ends the path and leaves the throughput
    bv0 = bv.copy(); bv0[:5] = 0
    t, _, _, alive = O.scatter_f64(0.5, wi, ones, None, pb, bv0, c, thp)
    assert not alive[:5].any() and np.array_equal(t[:5], thp[:5])


def _slice_scale(mx, R):
    """The grid-gradient fixed-point scale exponent (k_train.hip slice_scale):
    S = 62 - bits(R) - e with mx < 2^e (frexp) and R < 2^bits(R)."""
    if not mx > 0:
        return 0
    e = np.frexp(np.float32(mx))[1]
    lb = int(R).bit_length()
    return 62 - lb - int(e)


def test_fixed_point_grid_gradient_scale_cannot_overflow():
    """Worst case of the fixed-point LDS sums: R updates, all equal to the
    level's largest |update| mx, rounded to integers at scale 2^S, sum below
    2^63 (the int64 accumulator), while one update keeps at least 40 bits below
    mx for slices of up to 2^22 records."""
    rng = np.random.default_rng(7)
    for _ in range(2000):
        mx = float(np.float32(10.0 ** rng.uniform(-38, 38)))
        R = int(rng.integers(1, 1 << 31))
        S = _slice_scale(mx, R)
        q = int(np.rint(np.ldexp(np.float64(mx), S)))  # the largest single update
        assert R * q < (1 << 63), (mx, R, S)
        if R <= (1 << 22):
            assert np.ldexp(np.float64(mx), S) >= 2.0 ** 39  # resolution 2^-S <= mx * 2^-39
    # exact powers of two and a single record
    for mx in (2.0 ** -126, 1.0, 2.0 ** 100, np.float32(3.4028235e38)):
        S = _slice_scale(float(mx), 1)
        assert int(np.rint(np.ldexp(np.float64(mx), S))) < (1 << 63)
    assert _slice_scale(0.0, 5) == 0


# ----------------------------------------------------------------------------
# Hand-worked pins: NRC substitution on hand-built records, the DF loss at
# uniform logits, Adam and EMA under a constant gradient.
# ----------------------------------------------------------------------------
def _hand_records():
    """Three pixels, four depths, every value chosen so that no record is
    filtered (finite radiance, throughput above eps):
      pixel 0: cur_depth 4, no delta or miss vertex;
      pixel 1: cur_depth 4, vertex 1 delta, vertex 2 a miss;
      pixel 2: cur_depth 3, vertex 1 delta."""
    from gpu_util import random_records
    rec, aabb = random_records(3, max_depth=4, seed=90)
    D, N = 4, 3
    rec["cur_depth"] = np.array([4, 4, 3], np.uint32)
    rec["delta"] = np.zeros((D, N), np.uint8)
    rec["miss"] = np.zeros((D, N), np.uint8)
    rec["delta"][1, 1] = 1
    rec["miss"][2, 1] = 1
    rec["delta"][1, 2] = 1
    rec["L"] = (1.0 + np.arange(D * N * 3, dtype=np.float32).reshape(D, N, 3)) * np.float32(0.25)
    rec["thp"] = np.full((D, N, 3), 0.5, np.float32) + np.arange(D, dtype=np.float32)[:, None, None] * np.float32(0.125)
    rec["Le"] = np.full((D, N, 3), 0.0625, np.float32)
    rec["wi_pdf"] = np.full((D, N), 0.5, np.float32)
    return rec, aabb


def test_nrc_substitution_hand_worked():
    """OPTIMIZE_ON_NRC = 4 (train.cu:156-213, 265-308; workqueue.h:113-138) on
    hand-built records, the answers worked out by hand:
      inputs: every vertex below cur_depth that is neither delta nor a miss,
              pixel-major;  cur_depth drops by one;
      apply (replaceIndirectRadianceSkipDelta(p, k - 1, R_k)): the first
              non-delta vertex d <= k - 1 still below the new cur_depth gets
              L = thp[k-1] * (R_k + Le[k-1]);
      divide (divideRadiance(p, k, R_k)): L[k] /= R_k per channel where
              R_k > eps, for k below the new cur_depth."""
    rec, aabb = _hand_records()
    _, verts, _, depth = O.nrc_substitute_f64(rec, 3, 0, 1, aabb, None)
    assert [tuple(v) for v in verts] == [(0, 0), (1, 0), (2, 0), (3, 0), (0, 1), (3, 1), (0, 2), (2, 2)]
    assert list(depth) == [3, 3, 2]
    R = np.array([[2.0, 4.0, 1e-6]] * len(verts)) * (1 + np.arange(len(verts)))[:, None]
    R[:, 2] = [1e-6, 3.0, 1e-6, 5.0, 2.0, 1e-6, 7.0, 1.5]   # some channels at or below eps: not divided
    _, _, L, depth = O.nrc_substitute_f64(rec, 3, 0, 1, aabb, R)
    L0 = rec["L"].astype(np.float64)
    thp, Le = rec["thp"].astype(np.float64), rec["Le"].astype(np.float64)

    def div(v, r):
        return np.where(r > 1e-5, v / np.where(r > 1e-5, r, 1), v)

    exp = L0.copy()
    # pixel 0: R for vertices 0..3 are rows 0..3
    exp[0, 0] = div(thp[0, 0] * (R[1] + Le[0, 0]), R[0])
    exp[1, 0] = div(thp[1, 0] * (R[2] + Le[1, 0]), R[1])
    exp[2, 0] = div(thp[2, 0] * (R[3] + Le[2, 0]), R[2])
    # vertex 3 of pixel 0 is at the new cur_depth: untouched
    # pixel 1: inputs are vertices 0 (row 4) and 3 (row 5); vertex 3 writes
    # depth 2 (a miss, not a delta) from its own predecessor's thp and Le
    exp[2, 1] = thp[2, 1] * (R[5] + Le[2, 1])
    exp[0, 1] = div(L0[0, 1], R[4])
    # pixel 2: inputs are vertices 0 (row 6) and 2 (row 7); vertex 2's
    # predecessor 1 is a delta, so depth 0 gets thp[1] * (R + Le[1])
    exp[0, 2] = div(thp[1, 2] * (R[7] + Le[1, 2]), R[6])
    assert np.allclose(L, exp, rtol=1e-12, atol=0)
    assert list(depth) == [3, 3, 2]


def test_df_loss_closed_form_at_uniform_logits():
    """With the output layers zeroed every logit is 0, the bins are all 1 and
    q = p_theta p_phi + 1e-5 = 1 + 1e-5 for any direction, so
    loss = mean_i -Li / (wiPdf_i + 1e-5) * ln(1 + 1e-5), Li = mean of L_rgb
    (train.h:931-965) — with and without the fp16 bin storage."""
    cfg = G.hash_config(with_nrc=False, n_levels=4, log2_hashmap_size=8)
    th, ph, _ = O.config_layouts(cfg)
    rng = np.random.default_rng(95)
    pt, pp = rng.normal(0, 0.3, th["n_params"]), rng.normal(0, 0.3, ph["n_params"])
    for L, p in ((th, pt), (ph, pp)):
        p[L["layer_off"][L["n_hidden"]]:L["n_matrix"]] = 0.0
    from gpu_util import df_features, train_targets
    B = 64
    feat, tgt = df_features(B, 96), train_targets(B, 97)
    tgt[:, 5] = rng.uniform(0.05, 3.0, B)   # wiPdf per row
    eps = float(np.float32(1e-5))
    Li = tgt[:, 2:5].astype(np.float64).mean(1)
    expect = np.mean(-Li / (tgt[:, 5].astype(np.float64) + eps) * math.log(1 + eps))
    for half in (False, True):
        loss, _, _ = O.df_train_f64(th, ph, pt, pp, feat, 9, tgt, 1, grads=False, half_bins=half)
        assert abs(loss - expect) <= 1e-12 * abs(expect), (half, loss, expect)


def test_adam_constant_gradient_closed_form():
    """Under a constant gradient g the bias-corrected Adam step is
    lr * sqrt(1-b2^t)/(1-b1^t) * g(1-b1^t) / (|g| sqrt(1-b2^t) + eps), i.e.
    lr * sign(g) to eps / |g|: after t steps w = w0 - t lr sign(g).  A grid
    (non-matrix) parameter whose gradient is 0 on a step is skipped entirely
    (its own step counter stays), so (g, 0, g) moves it by 2 lr; a matrix
    parameter is updated on the zero step from its decayed moments."""
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-15
    n, n_matrix = 6, 3
    w0 = np.array([0.5, -0.25, 1.0, 0.75, -0.5, 2.0])
    g = np.array([0.3, -2.0, 1e-3, 0.7, -0.05, 5.0])
    m, v, st, w = np.zeros(n), np.zeros(n), np.zeros(n, np.uint32), w0.copy()
    for t in range(1, 5):
        w, m, v, st, _ = O.adam_ema_f64(n_matrix, lr, b1, b2, eps, 0.0, 0.95, t, g, w, m, v, st)
        assert np.allclose(w, w0 - t * lr * np.sign(g), rtol=0, atol=1e-12), t
        assert np.allclose(m, g * (1 - b1 ** t), rtol=1e-12) and np.allclose(v, g * g * (1 - b2 ** t), rtol=1e-12)
    # a zero gradient on step 2 of 3
    m, v, st, w = np.zeros(n), np.zeros(n), np.zeros(n, np.uint32), w0.copy()
    for t, gg in enumerate((g, np.zeros(n), g), 1):
        w, m, v, st, _ = O.adam_ema_f64(n_matrix, lr, b1, b2, eps, 0.0, 0.95, t, gg, w, m, v, st)
    assert list(st) == [3, 3, 3, 2, 2, 2]
    assert np.allclose(w[3:], w0[3:] - 2 * lr * np.sign(g[3:]), rtol=0, atol=1e-12)
    # matrix params moved on all three steps: lr on step 1; on step 2 by the
    # decayed moments m = b1 (1-b1) g, v = b2 (1-b2) g^2 at t = 2; on step 3
    # by m = (b1^2 (1-b1) + 1-b1) g, v = (b2^2 (1-b2) + 1-b2) g^2 at t = 3
    s2 = lr * math.sqrt(1 - b2 ** 2) / (1 - b1 ** 2) * b1 * (1 - b1) / math.sqrt(b2 * (1 - b2))
    s3 = lr * math.sqrt(1 - b2 ** 3) / (1 - b1 ** 3) * (b1 * b1 * (1 - b1) + 1 - b1) / math.sqrt(
        b2 * b2 * (1 - b2) + 1 - b2)
    assert np.allclose(w[:3], w0[:3] - (lr + s2 + s3) * np.sign(g[:3]), rtol=0, atol=1e-10)
    # l2 only on matrix params: with l2 > 0 and g = 0 a grid entry stays put
    w2, _, _, st2, _ = O.adam_ema_f64(n_matrix, lr, b1, b2, eps, 1e-2, 0.95, 1, np.zeros(n), w0, np.zeros(n),
                                      np.zeros(n), np.zeros(n, np.uint32))
    assert np.array_equal(w2[3:], w0[3:]) and list(st2) == [1, 1, 1, 0, 0, 0]
    assert np.allclose(w2[:3], w0[:3] - lr * np.sign(w0[:3]), rtol=0, atol=1e-12)


def test_ema_is_the_debiased_exponential_average():
    """EmaOptimizer (ema.h:62-136): after t steps of weights w_1..w_t the EMA
    is sum_s d^(t-s) (1-d) w_s / (1 - d^t), the normalised exponential average;
    the recursion's debias terms must reproduce that sum at every t."""
    d = 0.95
    rng = np.random.default_rng(98)
    n = 5
    ws = [rng.normal(0, 1, n) for _ in range(6)]
    ema = rng.normal(0, 1, n)   # the initial value has zero weight after debiasing
    for t in range(1, 7):
        _, _, _, _, ema = O.adam_ema_f64(0, 0.0, 0.9, 0.99, 1e-8, 0.0, d, t, np.zeros(n), ws[t - 1], np.zeros(n),
                                         np.zeros(n), np.zeros(n, np.uint32), w_lowp=ws[t - 1], ema=ema)
        expect = sum(d ** (t - s) * (1 - d) * ws[s - 1] for s in range(1, t + 1)) / (1 - d ** t)
        assert np.allclose(ema, expect, rtol=1e-12, atol=1e-12), t


# ----------------------------------------------------------------------------
# The consumer stage: sampleWi's decision and the record producers
# ----------------------------------------------------------------------------
def test_sphere_maps_closed_forms():
    """cartesianToSpherical and uniformSphereToSquare at the axes (closed
    forms), scale invariance, the zero vector (Eigen's normalized() leaves it)."""
    c2s, s2q = O.cartesian_to_spherical, O.uniform_sphere_to_square_one
    assert np.allclose(c2s([0, 0, 1]), [0, 0]) and np.allclose(c2s([0, 0, -2]), [math.pi, 0])
    assert np.allclose(c2s([1, 0, 0]), [math.pi / 2, 0]) and np.allclose(c2s([0, 1, 0]), [math.pi / 2, math.pi / 2])
    assert np.allclose(c2s([0, -3, 0]), [math.pi / 2, 1.5 * math.pi]) and np.allclose(c2s([-1, 0, 0]), [math.pi / 2, math.pi])
    assert np.allclose(c2s([0, 0, 0]), [math.pi / 2, 0])
    assert np.allclose(s2q([0, 0, 1]), [0, 0]) and np.allclose(s2q([0, 0, -1]), [1, 0])
    assert np.allclose(s2q([0, 1, 0]), [0.5, 0.25]) and np.allclose(s2q([0, -5, 0]), [0.5, 0.75])
    rng = np.random.default_rng(60)
    for v in rng.normal(0, 1, (50, 3)):
        sq = s2q(v)
        assert np.allclose(O.square_to_uniform_sphere(sq[None])[0], v / np.linalg.norm(v))   # inverse maps
        th, ph = c2s(v)
        assert np.allclose(O.spherical_to_cartesian(th, ph), v / np.linalg.norm(v))


def test_sample_wi_decisions_hand_worked():
    """sampleWi (integrator.cpp:355-401), row by row with rr 0.8, f 0.3,
    max guided depth 3:
      0: u_rr 0.9 >= 0.8 -> killed, wi and thp untouched
      1: smooth, depth 0, u_mix 0.5 >= 0.3 -> guided; thp / 0.8; wi[2:] = 1
      2: smooth, u_mix 0.1 < 0.3 -> BSDF sampled, mix; wi[:2] = square(+y)
      3: not smooth -> BSDF, no mix; its sample has f = 0: wi[:2] kept
      4: depth 3 (guiding off there) -> BSDF, no mix; square(-z) = (1, 0)
      5: smooth, u_mix 0.2 but bsdf pdf 0 -> BSDF, mix, wi[:2] kept"""
    depth = np.array([0, 0, 1, 0, 3, 2])
    smooth = np.array([1, 1, 1, 0, 1, 1], np.uint8)
    u_rr = np.array([0.9, 0.1, 0.5, 0.79, 0.0, 0.3])
    u_mix = np.array([0.5, 0.5, 0.1, 0.9, 0.9, 0.2])
    bwi = np.array([[1, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 1], [0, 0, -1], [1, 1, 1]], np.float64)
    bpdf = np.array([1, 1, 0.5, 2.0, 0.3, 0.0])
    bf = np.array([[1, 1, 1], [1, 1, 1], [0.2, 0, 0], [0, 0, 0], [0.5, 0.5, 0.5], [1, 1, 1]], np.float64)
    wi = np.full((6, 4), 0.125)
    thp = np.ones((6, 3))
    k, g, m, w, t = O.sample_wi_f64(0.8, 0.3, True, 3, depth, smooth, u_rr, u_mix, bwi, bpdf, bf, wi, thp)
    assert list(k) == [1, 0, 0, 0, 0, 0] and list(g) == [0, 1, 0, 0, 0, 0] and list(m) == [0, 0, 1, 0, 0, 1]
    assert np.array_equal(w[0], wi[0]) and np.array_equal(t[0], thp[0])
    assert np.allclose(t[1:], 1.25)
    assert np.allclose(w[1:, 2:], 1.0)
    assert np.allclose(w[1, :2], 0.125) and np.allclose(w[2, :2], [0.5, 0.25])
    assert np.allclose(w[3, :2], 0.125) and np.allclose(w[4, :2], [1.0, 0.0]) and np.allclose(w[5, :2], 0.125)
    # f = 0: always guided where guiding is on; f = 1: never, and no mixture
    _, g0, m0, _, _ = O.sample_wi_f64(0.8, 0.0, True, 3, depth, smooth, u_rr, u_mix, bwi, bpdf, bf, wi)
    _, g1, m1, _, _ = O.sample_wi_f64(0.8, 1.0, True, 3, depth, smooth, u_rr, u_mix, bwi, bpdf, bf, wi)
    assert list(g0) == [0, 1, 1, 0, 0, 1] and list(m0) == [0] * 6
    assert list(g1) == [0] * 6 and list(m1) == [0] * 6


def _empty_records(D, N):
    return {"cur_depth": np.zeros(N, np.int64), "pos": np.zeros((D, N, 3)), "dir": np.zeros((D, N, 2)),
            "wo": np.zeros((D, N, 2)), "normal": np.zeros((D, N, 2)), "roughness": np.zeros((D, N)),
            "L": np.zeros((D, N, 3)), "thp": np.zeros((D, N, 3)), "bsdf_val": np.zeros((D, N, 3)),
            "diffuse": np.zeros((D, N, 3)), "specular": np.zeros((D, N, 3)), "wi_pdf": np.zeros((D, N)),
            "mis_weight": np.zeros((D, N)), "delta": np.zeros((D, N), np.uint8), "miss": np.zeros((D, N), np.uint8),
            "Le": np.zeros((D, N, 3))}


def test_record_producers_hand_worked():
    """incrementDepth / incrementDepthLastBounceNRC / recordRadiance /
    overwriteLe (workqueue.h:26-106) on a 2-deep buffer of 4 pixels, training
    pixels 1 and 3 (offset 1, stride 2)."""
    rec = _empty_records(2, 4)
    sd = {"pos": np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float64),
          "normal": np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], np.float64),
          "wo": np.array([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], np.float64),
          "roughness": np.array([0.1, 0.2, 0.3]), "diffuse": np.full((3, 3), 0.4), "specular": np.full((3, 3), 0.6),
          "smooth": np.array([1, 0, 1], np.uint8)}
    pixel, depth = np.array([1, 3, 2]), np.array([0, 0, 0])
    alive = np.array([1, 0, 1], np.uint8)
    wi_dir = np.array([[0, 0, 3], [1, 1, 1], [1, 0, 0]], np.float64)
    thp = np.array([[0.5, 0.25, 1.0], [9, 9, 9], [9, 9, 9]])
    O.record_scatter_f64(rec, sd, pixel, depth, True, 5, 1, 2, True, True, alive=alive, wi_dir=wi_dir, thp=thp,
                         scatter_pdf=np.array([0.7, 0, 0.1]), mis=np.array([0.6, 0, 0]),
                         bsdf_val=np.array([[0.2, 0.4, 0.8], [0, 0, 0], [1, 1, 1]]), cos_theta=np.array([-0.5, 0, 1]),
                         delta=np.array([1, 0, 0], np.uint8))
    # pixel 1 scattered: every field of vertex 0
    assert list(rec["cur_depth"]) == [0, 1, 0, 1]
    assert np.allclose(rec["thp"][0, 1], [0.5, 0.25, 1.0]) and np.allclose(rec["pos"][0, 1], [1, 2, 3])
    assert np.allclose(rec["dir"][0, 1], [0, 0]) and np.allclose(rec["normal"][0, 1], [0, 0])
    assert np.allclose(rec["wo"][0, 1], [math.pi / 2, math.pi / 2]) and rec["wi_pdf"][0, 1] == 0.7
    assert rec["mis_weight"][0, 1] == 0.6 and rec["delta"][0, 1] == 1 and rec["miss"][0, 1] == 0
    assert np.allclose(rec["bsdf_val"][0, 1], [0.1, 0.2, 0.4])     # product with |cos| = 0.5
    assert rec["roughness"][0, 1] == 0.1 and np.allclose(rec["diffuse"][0, 1], 0.4)
    # pixel 3 ended (not alive): the placeholder from its shading data; not smooth -> delta
    assert np.allclose(rec["pos"][0, 3], [4, 5, 6]) and rec["delta"][0, 3] == 1 and rec["miss"][0, 3] == 0
    assert np.allclose(rec["normal"][0, 3], [math.pi / 2, 0]) and np.allclose(rec["wo"][0, 3], [math.pi, 0])
    assert np.allclose(rec["thp"][0, 3], 0) and rec["wi_pdf"][0, 3] == 0     # dir / pdf / mis left as they were
    # pixel 2 is not a training pixel: untouched
    assert rec["cur_depth"][2] == 0 and np.allclose(rec["pos"][:, 2], 0)
    # mis_aware off -> misWeight 1; a third vertex does not fit the 2-deep buffer
    O.record_scatter_f64(rec, sd, np.array([1]), np.array([1]), True, 5, 1, 2, False, False, alive=np.array([1]),
                         wi_dir=wi_dir[2:], thp=np.array([[2.0, 2.0, 2.0]]), scatter_pdf=np.array([0.9]),
                         mis=np.array([0.3]), bsdf_val=np.array([[1.0, 0.5, 0.25]]), cos_theta=np.array([0.5]))
    assert rec["cur_depth"][1] == 2 and rec["mis_weight"][1, 1] == 1.0 and np.allclose(rec["bsdf_val"][1, 1], [1, 0.5, 0.25])
    assert np.allclose(rec["dir"][1, 1], [math.pi / 2, 0])   # +x
    before = {k: v.copy() for k, v in rec.items()}
    O.record_scatter_f64(rec, sd, np.array([1]), np.array([1]), True, 5, 1, 2, True, False, alive=np.array([1]),
                         wi_dir=wi_dir[:1], thp=thp[:1], scatter_pdf=np.array([1.0]), mis=np.array([1.0]),
                         bsdf_val=np.array([[1.0, 1, 1]]), cos_theta=np.array([1.0]))
    assert all(np.array_equal(before[k], rec[k]) for k in rec)
    # depth at max_train_depth: no record
    O.record_scatter_f64(rec, sd, np.array([3]), np.array([5]), True, 5, 1, 2, True, False)
    assert rec["cur_depth"][3] == 1
    # radiance: added to every vertex below cur_depth, Le onto the last one
    O.record_radiance_f64(rec, np.array([1, 3, 2]), np.array([[1.0, 2, 3], [0.5, 0.5, 0.5], [7, 7, 7]]), True, 5, 1, 2,
                          Le=np.array([[4.0, 4, 4], [8, 8, 8], [1, 1, 1]]))
    assert np.allclose(rec["L"][0, 1], [1, 2, 3]) and np.allclose(rec["L"][1, 1], [1, 2, 3])
    assert np.allclose(rec["Le"][1, 1], 4) and np.allclose(rec["Le"][0, 1], 0)
    assert np.allclose(rec["L"][0, 3], 0.5) and np.allclose(rec["Le"][0, 3], 8) and np.allclose(rec["L"][:, 2], 0)
    # a miss at depth <= max_train_depth adds the placeholder (empty ShadingData)
    O.record_radiance_f64(rec, np.array([3]), np.array([[0.0, 0, 0]]), True, 1, 1, 2, miss=np.array([1]),
                          depth=np.array([1]))
    assert rec["cur_depth"][3] == 2 and rec["miss"][1, 3] == 1 and rec["delta"][1, 3] == 0
    assert rec["roughness"][1, 3] == 1.0 and np.allclose(rec["pos"][1, 3], 0)
    O.record_radiance_f64(rec, np.array([3]), np.array([[0.0, 0, 0]]), True, 1, 1, 2, miss=np.array([1]),
                          depth=np.array([2]))
    assert rec["cur_depth"][3] == 2


def test_recorded_path_yields_arriving_radiance():
    """A simulated path: vertex k scatters with throughput T_k (product of
    the per-bounce factors), then an emitter of radiance E is hit.  The
    producers store L_k = E T_path at every vertex, so the DF training target
    L_k / T_k * mis * bsdfVal (train.cu:6-96) is the radiance arriving at
    vertex k, E prod_{j>k} factor_j, times bsdfVal and mis."""
    D, N = 4, 1
    rec = _empty_records(D, N)
    rng = np.random.default_rng(61)
    factors = rng.uniform(0.3, 0.9, (3, 3))
    T = np.ones(3)
    sd = {"pos": rng.uniform(-1, 1, (1, 3)), "normal": np.array([[0.0, 0, 1]]), "wo": np.array([[0.0, 1, 1]]),
          "roughness": np.array([0.5]), "diffuse": np.full((1, 3), 0.5), "specular": np.zeros((1, 3)),
          "smooth": np.array([1], np.uint8)}
    for k in range(3):
        T = T * factors[k]
        O.record_scatter_f64(rec, sd, np.array([0]), np.array([k]), True, 5, 0, 1, True, False, alive=np.array([1]),
                             wi_dir=rng.normal(0, 1, (1, 3)), thp=T[None], scatter_pdf=np.array([0.4]),
                             mis=np.array([0.75]), bsdf_val=np.array([[0.3, 0.3, 0.3]]), cos_theta=np.array([0.5]))
    E = np.array([2.0, 1.0, 0.5])
    O.record_radiance_f64(rec, np.array([0]), (E * T)[None], True, 5, 0, 1, Le=E[None])
    rec["cur_depth"] = rec["cur_depth"].astype(np.uint32)
    aabb = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    f, t = O.training_targets_f64(rec, 1, 0, 1, aabb)
    assert f.shape == (3, 9)
    for k in range(3):
        arriving = E * np.prod(factors[k + 1:], axis=0)
        assert np.allclose(t[k, 2:5], arriving * 0.75 * 0.3, rtol=1e-12), k
        assert t[k, 5] == 0.4


def test_underflowing_softmax_rows_give_a_nan_loss_and_no_gradient():
    """theta logits all near -1e6: the softmax whose maximum starts at 0
    (train.h:846) sums to exactly zero, its bins and pdf are not finite, so the
    batch loss is NaN; the loss-head gradient of such a row counts as zero, so
    the whole gradient is zero.  With one finite row (theta's last layer left
    as drawn) the gradient is finite and nonzero again."""
    cfg = G.hash_config(with_nrc=False, log2_hashmap_size=12)
    Lt, Lp, _ = O.config_layouts(cfg)
    rng = np.random.default_rng(5)
    pt, pp = trained_scale_params(Lt, rng), trained_scale_params(Lp, rng)
    off = Lt["layer_off"]
    bad = pt.copy()
    bad[off[1]:off[3]] = np.abs(bad[off[1]:off[3]])
    bad[off[3]:off[4]] = -1000.0
    B = 64
    feat, tgt = df_features(B, 150), train_targets(B, 151)
    loss, gt, gp = O.df_train_f64(Lt, Lp, bad, pp, feat, 9, tgt, True)
    assert math.isnan(loss)
    assert not np.any(gt) and not np.any(gp)
    loss, gt, gp = O.df_train_f64(Lt, Lp, pt, pp, feat, 9, tgt, True)
    assert math.isfinite(loss) and np.all(np.isfinite(gt)) and np.any(gt) and np.any(gp)
