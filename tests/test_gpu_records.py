"""GPU tests of training from recorded path vertices: row generation against
the FP64 oracle restatement of generate_training_data_AR / _nrc, and the
per-frame training passes against the same steps run by hand."""
import numpy as np
import pytest

import dfguide as G
import oracle as O
from gpu_util import random_records, dev, host

pytestmark = pytest.mark.gpu


def _dev_records(torch, rec):
    out = {}
    for k, v in rec.items():
        if k == "cur_depth":
            out[k] = torch.from_numpy(v.view(np.int32)).cuda()
        else:
            out[k] = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    return out


@pytest.mark.parametrize("n_pixels,offset,stride", [(3000, 0, 1), (3001, 2, 3), (64, 0, 1)])
def test_training_rows_match_oracle(torch_cuda, n_pixels, offset, stride):
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=True, log2_hashmap_size=12))
    rec, aabb = random_records(n_pixels, seed=n_pixels + offset)
    drec, daabb = _dev_records(torch, rec), dev(torch, aabb)
    n_train = (n_pixels - offset + stride - 1) // stride
    fo, to = O.training_targets_f64(rec, n_train, offset, stride, aabb)
    cap = n_train * 5
    feat = torch.zeros((cap, 9), device="cuda")
    tgt = torch.zeros((cap, 9), device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    f.training_targets(drec, n_train, daabb, feat, tgt, cnt, offset=offset, stride=stride)
    torch.cuda.synchronize()
    m = int(host(cnt)[0])
    assert m == len(fo), (m, len(fo))
    assert np.allclose(host(feat)[:m], fo, rtol=1e-5, atol=2e-6)
    assert np.allclose(host(tgt)[:m], to, rtol=1e-5, atol=2e-6)
    # NRC rows
    fno, tno = O.nrc_training_targets_f64(rec, n_train, offset, stride, aabb)
    featn = torch.zeros((cap, 15), device="cuda")
    tgtn = torch.zeros((cap, 3), device="cuda")
    f.nrc_training_targets(drec, n_train, daabb, featn, tgtn, cnt, offset=offset, stride=stride)
    torch.cuda.synchronize()
    m = int(host(cnt)[0])
    assert m == len(fno), (m, len(fno))
    assert np.allclose(host(featn)[:m], fno, rtol=1e-5, atol=2e-6)
    assert np.allclose(host(tgtn)[:m], tno, rtol=1e-5, atol=1e-6)
    f.close()


def test_training_rows_capacity_and_errors(torch_cuda):
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False, log2_hashmap_size=12))
    rec, aabb = random_records(500, seed=5)
    drec, daabb = _dev_records(torch, rec), dev(torch, aabb)
    fo, to = O.training_targets_f64(rec, 500, 0, 1, aabb)
    cap = len(fo) // 2
    feat = torch.full((cap, 9), -7.0, device="cuda")
    tgt = torch.full((cap, 9), -7.0, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    f.training_targets(drec, 500, daabb, feat, tgt, cnt)
    torch.cuda.synchronize()
    assert int(host(cnt)[0]) == len(fo)            # the count covers every valid row
    assert np.allclose(host(feat), fo[:cap], rtol=1e-5, atol=2e-6)
    with pytest.raises(G.DfgError):                # training pixels beyond the buffer
        f.training_targets(drec, 501, daabb, feat, tgt, cnt)
    with pytest.raises(G.DfgError):
        f.training_targets(drec, 200, daabb, feat, tgt, cnt, offset=10, stride=3)
    f.training_targets(drec, 0, daabb, feat, tgt, cnt)
    torch.cuda.synchronize()
    assert int(host(cnt)[0]) == 0
    f.close()


def test_train_records_runs_the_reference_schedule(torch_cuda):
    """dfg_train_records = row generation + the learnedDistribution.cpp batch
    loop; the same slices trained by hand on an identical field give the same
    weights, bit for bit."""
    torch = torch_cuda
    cfg = G.hash_config(with_nrc=False, log2_hashmap_size=14)
    fa, fb = G.GuidingField(cfg), G.GuidingField(cfg)
    rec, aabb = random_records(100000, seed=9)
    drec, daabb = _dev_records(torch, rec), dev(torch, aabb)
    bs, bpf = 65536, 5
    rows, steps, loss = fa.train_records(drec, 100000, daabb, batch_size=bs, batch_per_frame=bpf)
    fo, to = O.training_targets_f64(rec, 100000, 0, 1, aabb)
    sched, planned = O.df_batches(len(fo), bs, bpf)
    assert rows == len(fo) and steps == len(sched) and steps >= 2
    feat = torch.zeros((100000 * 5, 9), device="cuda")
    tgt = torch.zeros((100000 * 5, 9), device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    fb.training_targets(drec, 100000, daabb, feat, tgt, cnt)
    losses = torch.zeros(len(sched), device="cuda")
    for i, (r0, b) in enumerate(sched):
        fb.train_step(feat[r0:r0 + b].contiguous(), tgt[r0:r0 + b].contiguous(), losses[i:i + 1])
    torch.cuda.synchronize()
    assert abs(loss - float(host(losses).sum()) / planned) <= 1e-4 * abs(loss) + 1e-6
    # the training step is deterministic (fixed-point grid-gradient sums,
    # weight-gradient partial sums added in a fixed order), so the records
    # pass and the hand-run steps end on identical weights
    for k in (0, 1):
        for which in (G.DFG_PARAMS_MASTER, G.DFG_PARAMS_INFERENCE):
            a, b = fa.get_params(k, which), fb.get_params(k, which)
            assert np.array_equal(a, b), "net %d view %d: %d weights differ" % (k, which, (a != b).sum())
    fa.close()
    fb.close()


def test_nrc_train_records_schedule(torch_cuda):
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=True, log2_hashmap_size=12))
    rec, aabb = random_records(150000, seed=11)
    drec, daabb = _dev_records(torch, rec), dev(torch, aabb)
    w0 = f.get_params(2, G.DFG_PARAMS_MASTER)
    rows, steps, loss = f.nrc_train_records(drec, 150000, daabb, batch_size=131072, batch_per_frame=2)
    fo, _ = O.nrc_training_targets_f64(rec, 150000, 0, 1, aabb)
    sched, planned = O.nrc_batches(len(fo), 131072, 2)
    assert rows == len(fo) and steps == len(sched) and steps >= 1
    assert np.isfinite(loss) and loss > 0 and not np.array_equal(w0, f.get_params(2, G.DFG_PARAMS_MASTER))
    # a schedule whose slices fall under the 65,536-row minimum trains nothing
    rows, steps, loss = f.nrc_train_records(drec, 150000, daabb, batch_size=65536, batch_per_frame=4)
    sched, _ = O.nrc_batches(len(fo), 65536, 4)
    assert steps == len(sched) == 0 and loss == 0.0
    f.close()


def test_nrc_substitution_matches_oracle(torch_cuda):
    """Radiance-cache substitution of the records (the reference's
    OPTIMIZE_ON_NRC = 4 path): new cur_depth exact; new L against the oracle's
    apply / divide rules fed with the cache's output for the same inputs
    (computed from the oracle's inputs, so within the cache's fp16 input
    rounding of the library's own)."""
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=True, log2_hashmap_size=12))
    rec, aabb = random_records(2000, seed=21)
    rec["Le"] = np.random.default_rng(22).uniform(0, 0.5, rec["thp"].shape).astype(np.float32)
    rec["cur_depth"] = np.minimum(rec["cur_depth"], 4)
    rec["cur_depth"][:7] = 0                       # pixels without vertices stay at 0
    drec, daabb = _dev_records(torch, rec), dev(torch, aabb)
    fo, vk, _, _ = O.nrc_substitute_f64(rec, 1999, 1, 1, aabb, None)
    # the library's own cache inputs for those vertices: the NRC training rows
    # of the records with one more vertex per pixel are the same set, in order
    plus = dict(drec)
    plus["cur_depth"] = dev(torch, (rec["cur_depth"] + 1).view(np.int32))
    featn = torch.zeros((1999 * 5, 15), device="cuda")
    tgtn = torch.zeros((1999 * 5, 3), device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    f.nrc_training_targets(plus, 1999, daabb, featn, tgtn, cnt, offset=1, stride=1)
    torch.cuda.synchronize()
    assert int(host(cnt)[0]) == len(vk)
    assert np.allclose(host(featn)[:len(vk)], fo, rtol=1e-5, atol=2e-6)
    rgb = torch.zeros((len(vk), 3), device="cuda")
    f.nrc_infer(featn[:len(vk)].contiguous(), rgb)
    m = f.nrc_substitute_radiance(drec, 1999, daabb, offset=1, stride=1)
    torch.cuda.synchronize()
    assert m == len(vk)
    _, _, L_o, depth_o = O.nrc_substitute_f64(rec, 1999, 1, 1, aabb, host(rgb).astype(np.float64))
    depth_g = host(drec["cur_depth"]).view(np.uint32)
    assert np.array_equal(depth_g.astype(np.int64), depth_o)
    assert depth_g[0] == rec["cur_depth"][0]        # outside the training pixels: untouched
    L_g = host(drec["L"])
    L0 = rec["L"].astype(np.float64)
    same = (L_o == L0) | (np.isnan(L_o) & np.isnan(L0))
    assert np.array_equal(L_g[same], rec["L"][same], equal_nan=True)
    assert np.allclose(L_g[~same], L_o[~same], rtol=1e-5, atol=1e-7, equal_nan=True)
    assert (~same).sum() > 100
    f.close()


def test_directions_and_scatter_match_oracle(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(70)
    n = 5000
    wi = np.concatenate([rng.random((n, 2)), rng.uniform(0.1, 3, (n, 2))], 1).astype(np.float32)
    d_wi = dev(torch, wi)
    d_dir = torch.zeros((n, 3), device="cuda")
    d_pdf = torch.zeros(n, device="cuda")
    G.directions(d_wi, d_dir, d_pdf)
    torch.cuda.synchronize()
    v = O.square_to_uniform_sphere(wi[:, :2].astype(np.float64))
    assert np.allclose(host(d_dir), v, atol=2e-6)
    assert np.allclose(O.uniform_sphere_to_square(host(d_dir).astype(np.float64))[:, 0], wi[:, 0], atol=1e-5)
    assert np.allclose(host(d_pdf), wi[:, 2] * wi[:, 3] / (4 * np.pi), rtol=1e-6)
    pb = rng.uniform(0.0, 2, n).astype(np.float32)
    pb[:50] = 0.0
    bv = rng.random((n, 3)).astype(np.float32)
    bv[50:100] = 0.0
    c = rng.uniform(-1, 1, n).astype(np.float32)
    thp = rng.random((n, 3)).astype(np.float32)
    guided = (rng.random(n) < 0.5).astype(np.uint8)
    mix = (rng.random(n) < 0.5).astype(np.uint8)
    d_thp = dev(torch, thp)
    d_p, d_m = torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    d_alive = torch.zeros(n, dtype=torch.uint8, device="cuda")
    G.scatter(d_wi, dev(torch, pb), dev(torch, bv), dev(torch, c), d_thp, bsdf_fraction=0.3, guided=dev(torch, guided),
              mix=dev(torch, mix), scatter_pdf=d_p, mis=d_m, alive=d_alive)
    torch.cuda.synchronize()
    t, p, mis, alive = O.scatter_f64(0.3, wi, guided, mix, pb, bv, c, thp)
    assert np.array_equal(host(d_alive) != 0, alive)
    assert np.allclose(host(d_p), p, rtol=2e-6, atol=1e-9)
    assert np.allclose(host(d_thp), t, rtol=2e-5, atol=1e-7)
    assert np.allclose(host(d_m)[alive], mis[alive], rtol=2e-5)


def test_pdf_image_matches_oracle_and_integrates_to_one(torch_cuda):
    torch = torch_cuda
    cfg = G.hash_config(with_nrc=False)
    f = G.GuidingField(cfg)
    th, ph, _ = O.config_layouts(cfg)
    from gpu_util import df_features
    feat = df_features(1, 80)[0]
    rt, rp = 128, 96
    img = torch.zeros((rt, rp), device="cuda")
    f.pdf_image(dev(torch, feat), img)
    torch.cuda.synchronize()
    u0 = (np.arange(rt) + 0.5) / rt
    u1 = (np.arange(rp) + 0.5) / rp
    wi = np.zeros((rt * rp, 4), np.float32)
    wi[:, 0] = np.repeat(u0, rp)
    wi[:, 1] = np.tile(u1, rt)
    wi[:, 2:] = 1.0
    o = O.query_f64(th, ph, f.get_params(0, G.DFG_PARAMS_INFERENCE), f.get_params(1, G.DFG_PARAMS_INFERENCE),
                    np.repeat(feat[None], rt * rp, 0), 9, np.zeros((rt * rp, 4), np.float32), wi, 1,
                    guided=np.zeros(rt * rp, np.uint8))
    ref = (np.abs(o[:, 2].astype(np.float64)) * np.abs(o[:, 3]) / (4 * np.pi)).reshape(rt, rp)
    g = host(img).astype(np.float64)
    rel = np.abs(g - ref) / ref
    assert np.median(rel) <= 1e-5 and rel.max() <= 1e-3
    # the density integrates to one over the sphere (midpoint rule on the square, 4 pi sr)
    assert abs(g.sum() * 4 * np.pi / (rt * rp) - 1.0) < 1e-2
    with pytest.raises(G.DfgError):
        f.pdf_image(dev(torch, feat), torch.zeros((0, 4), device="cuda"))
    f.close()


# ----------------------------------------------------------------------------
# The consumer stage: sampleWi's decision and the record producers
# ----------------------------------------------------------------------------
def _queue(rng, n, n_pixels):
    pixel = rng.permutation(n_pixels)[:n].astype(np.int32)
    pixel[:3] = [-1, n_pixels, n_pixels + 7]     # outside the records: ignored
    v = lambda *s: rng.normal(0, 1, s).astype(np.float32)
    sd = {"pos": v(n, 3), "normal": v(n, 3), "wo": v(n, 3), "roughness": rng.random(n).astype(np.float32),
          "diffuse": rng.random((n, 3)).astype(np.float32), "specular": rng.random((n, 3)).astype(np.float32),
          "smooth": (rng.random(n) < 0.8).astype(np.uint8)}
    sd["normal"][5] = 0.0                        # a zero vector: normalized() leaves it
    return pixel, rng.integers(0, 7, n).astype(np.int32), sd


def _rec_host(drec):
    out = {k: host(v) for k, v in drec.items()}
    out["cur_depth"] = out["cur_depth"].view(np.uint32).astype(np.int64)
    return out


def _rec_oracle(rec):
    out = {k: np.array(v, dtype=np.float64 if v.dtype.kind == "f" else v.dtype) for k, v in rec.items()}
    out["cur_depth"] = rec["cur_depth"].astype(np.int64)
    return out


def _assert_records_match(g, o):
    for k in o:
        if k in ("dir", "normal", "wo"):
            # spherical angles: acos of a normalised fp32 z near +-1 loses
            # digits (d theta ~ dz / sqrt(1 - z^2)); typical rows agree to 1e-6
            err = np.abs(g[k] - o[k])
            assert np.nanmax(err) <= 5e-5 and np.nanmedian(err) <= 1e-6, "record field %s differs" % k
        elif o[k].dtype.kind == "f":
            assert np.allclose(g[k], o[k], rtol=2e-6, atol=2e-6, equal_nan=True), "record field %s differs" % k
        else:
            assert np.array_equal(g[k].astype(np.int64), o[k].astype(np.int64)), "record field %s differs" % k


def test_sample_wi_matches_oracle(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(70)
    n = 5000
    depth = rng.integers(0, 7, n).astype(np.int32)
    smooth = (rng.random(n) < 0.8).astype(np.uint8)
    u_rr, u_mix = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    bwi = rng.normal(0, 1, (n, 3)).astype(np.float32)
    bwi[7] = 0.0
    bpdf = rng.random(n).astype(np.float32)
    bpdf[rng.random(n) < 0.05] = 0.0
    bf = rng.random((n, 3)).astype(np.float32)
    bf[rng.random(n) < 0.05] = 0.0
    wi = rng.random((n, 4)).astype(np.float32)
    thp = rng.random((n, 3)).astype(np.float32)
    for f in (0.0, 0.3, 1.0):
        d = {k: dev(torch, a) for k, a in (("depth", depth), ("smooth", smooth), ("u_rr", u_rr), ("u_mix", u_mix),
                                             ("bwi", bwi), ("bpdf", bpdf), ("bf", bf), ("wi", wi), ("thp", thp))}
        k_, g_, m_ = (torch.full((n,), 9, dtype=torch.uint8, device="cuda") for _ in range(3))
        G.sample_wi(d["depth"], d["smooth"], d["u_rr"], d["u_mix"], d["bwi"], d["bpdf"], d["bf"], d["wi"], rr_killed=k_,
                    guided=g_, mix=m_, thp=d["thp"], rr_prob=0.8, bsdf_fraction=f, guiding_enabled=True,
                    max_guided_depth=4)
        torch.cuda.synchronize()
        ko, go, mo, wo, to = O.sample_wi_f64(np.float32(0.8), np.float32(f), True, 4, depth, smooth, u_rr, u_mix, bwi,
                                             bpdf, bf, wi, thp)
        assert np.array_equal(host(k_), ko) and np.array_equal(host(g_), go) and np.array_equal(host(m_), mo)
        assert np.allclose(host(d["wi"]), wo, rtol=0, atol=2e-6)
        assert np.allclose(host(d["thp"]), to, rtol=1e-6, atol=0)
        assert 0 < ko.sum() and (go.sum() == 0 if f == 1.0 else 0 < go.sum() < n)


@pytest.mark.parametrize("offset,stride", [(0, 1), (3, 4)])
def test_record_producers_match_oracle(torch_cuda, offset, stride):
    """dfg_record_scatter (scattered and ended rows, RR-killed rows skipped)
    then dfg_record_radiance (emission and misses with placeholders) on
    records already holding a few vertices per pixel: every field against
    the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(71 + stride)
    n_pix = 3000
    rec, _ = random_records(n_pix, max_depth=5, seed=72)
    rec["cur_depth"] = rng.integers(0, 6, n_pix).astype(np.uint32)
    rec["Le"] = rng.random((5, n_pix, 3)).astype(np.float32)
    drec = _dev_records(torch, rec)
    orec = _rec_oracle(rec)
    n = 2000
    pixel, depth, sd = _queue(rng, n, n_pix)
    alive = (rng.random(n) < 0.8).astype(np.uint8)
    killed = (rng.random(n) < 0.1).astype(np.uint8)
    wi_dir = rng.normal(0, 1, (n, 3)).astype(np.float32)
    thp = rng.random((n, 3)).astype(np.float32)
    spdf, mis = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    bval, cos = rng.random((n, 3)).astype(np.float32), rng.uniform(-1, 1, n).astype(np.float32)
    delta = (rng.random(n) < 0.1).astype(np.uint8)
    dsd = {k: dev(torch, v) for k, v in sd.items()}
    G.record_scatter(drec, dsd, dev(torch, pixel), dev(torch, depth), alive=dev(torch, alive), wi_dir=dev(torch, wi_dir),
                     thp=dev(torch, thp), scatter_pdf=dev(torch, spdf), mis=dev(torch, mis), bsdf_val=dev(torch, bval),
                     cos_theta=dev(torch, cos), delta=dev(torch, delta), rr_killed=dev(torch, killed),
                     train_enabled=True, max_train_depth=4, pixel_offset=offset, pixel_stride=stride, mis_aware=True,
                     product_with_cosine=True)
    O.record_scatter_f64(orec, sd, pixel, depth, True, 4, offset, stride, True, True, rr_killed=killed, alive=alive,
                         wi_dir=wi_dir, thp=thp, scatter_pdf=spdf, mis=mis, bsdf_val=bval, cos_theta=cos, delta=delta)
    torch.cuda.synchronize()
    _assert_records_match(_rec_host(drec), orec)
    # the last bounce: every row an ended vertex (alive NULL)
    G.record_scatter(drec, dsd, dev(torch, pixel), dev(torch, depth), train_enabled=True, max_train_depth=6,
                     pixel_offset=offset, pixel_stride=stride)
    O.record_scatter_f64(orec, sd, pixel, depth, True, 6, offset, stride, True, False)
    torch.cuda.synchronize()
    _assert_records_match(_rec_host(drec), orec)
    # radiance of emitter hits and misses
    pixel2, depth2, _ = _queue(rng, n, n_pix)
    contrib, Le = rng.random((n, 3)).astype(np.float32), rng.random((n, 3)).astype(np.float32)
    miss = (rng.random(n) < 0.4).astype(np.uint8)
    G.record_radiance(drec, dev(torch, pixel2), dev(torch, contrib), Le=dev(torch, Le), miss=dev(torch, miss),
                      depth=dev(torch, depth2), train_enabled=True, max_train_depth=3, pixel_offset=offset,
                      pixel_stride=stride)
    O.record_radiance_f64(orec, pixel2, contrib, True, 3, offset, stride, Le=Le, miss=miss, depth=depth2)
    torch.cuda.synchronize()
    g = _rec_host(drec)
    _assert_records_match(g, orec)
    assert (g["cur_depth"] != rec["cur_depth"]).sum() > 100
    # training off: nothing changes
    before = {k: v.clone() for k, v in drec.items()}
    G.record_radiance(drec, dev(torch, pixel2), dev(torch, contrib), train_enabled=False)
    torch.cuda.synchronize()
    def same(a, b):   # the records hold NaN radiance on purpose
        if a.is_floating_point():
            return torch.equal(a.isnan(), b.isnan()) and torch.equal(a.nan_to_num(0.0), b.nan_to_num(0.0))
        return torch.equal(a, b)
    assert all(same(before[k], drec[k]) for k in drec)
