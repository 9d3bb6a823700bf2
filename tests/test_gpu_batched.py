"""GPU tests of the batched paths: wavefront frames (all bounces in one
persistent, graph-captured launch) and the continuous-batching server.
Every row must equal the same row queried alone through dfg_query."""
import numpy as np
import pytest

import dfguide as G
from gpu_util import df_features, uniforms, wi_init, train_targets, dev, host

pytestmark = pytest.mark.gpu


def _bounces(first, k, survival=0.7):
    sizes = [first]
    for _ in range(k - 1):
        sizes.append(int(sizes[-1] * survival))
    return sizes


def _alone(torch, f, feat, u, wi0, sizes, guided=None):
    out, row = [], 0
    for n in sizes:
        w = dev(torch, wi0[row:row + n])
        g = dev(torch, guided[row:row + n]) if guided is not None else None
        if n:
            f.query(dev(torch, feat[row:row + n]), dev(torch, u[row:row + n]), w, guided=g)
        out.append(w)
        row += n
    torch.cuda.synchronize()
    return np.concatenate([host(w) for w in out], 0) if out else np.zeros((0, 4), np.float32)


@pytest.mark.parametrize("tuning", [(4, 1, 0), (8, 2, 1), (2, 4, 1)])
def test_wavefront_matches_per_bounce_queries(torch_cuda, tuning):
    torch = torch_cuda
    cfg = G.hash_config(n_bins_theta=64, n_bins_phi=64, with_nrc=False)
    cfg["dfg"]["grid_prepass_min_rows"] = 1 << 30  # per-bounce reference: in-kernel gathers
    f = G.GuidingField(cfg)
    sizes = _bounces(20000, 6)
    cap = sum(sizes)
    feat, u = df_features(cap, 80), uniforms(cap, 81)
    guided = (np.random.default_rng(82).random(cap) < 0.8).astype(np.uint8)
    wi0 = wi_init(cap)
    wi0[:, :2] = np.random.default_rng(83).random((cap, 2), dtype=np.float32) * np.float32(0.999)
    ref = _alone(torch, f, feat, u, wi0, sizes, guided)
    wf = G.Wavefront(f, 8, cap)
    wf.set_tuning(*tuning)
    d_sizes = torch.tensor(sizes, dtype=torch.int64, device="cuda")
    d_feat, d_u, d_g = dev(torch, feat), dev(torch, u), dev(torch, guided)
    d_wi = dev(torch, wi0)
    wf.run(d_sizes, d_feat, d_u, d_wi, guided=d_g)
    torch.cuda.synchronize()
    assert np.array_equal(host(d_wi), ref)
    # replay the captured graph with new (smaller, uneven) sizes read on the device
    sizes2 = [7000, 0, 3333, 64, 1, 5000]
    n2 = sum(sizes2)
    ref2 = _alone(torch, f, feat[:n2], u[:n2], wi0[:n2], sizes2, guided[:n2])
    d_sizes.copy_(torch.tensor(sizes2, dtype=torch.int64, device="cuda"))
    d_wi.copy_(dev(torch, wi0))
    wf.run(d_sizes, d_feat, d_u, d_wi, guided=d_g)
    torch.cuda.synchronize()
    assert np.array_equal(host(d_wi)[:n2], ref2)
    assert np.array_equal(host(d_wi)[n2:], wi0[n2:])   # rows beyond the frame untouched
    wf.close()
    f.close()


def test_wavefront_graph_replays_with_random_sizes(torch_cuda):
    """One captured frame graph replayed 12 times with random bounce sizes
    (empty bounces, sizes that are not multiples of 64, 1 to 8 bounces) read
    on the device: every replay equals the per-bounce queries, and rows past
    the frame stay untouched."""
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False, log2_hashmap_size=14))
    cap = 60000
    feat, u = df_features(cap, 90), uniforms(cap, 91)
    wi0 = wi_init(cap)
    wf = G.Wavefront(f, 8, cap)
    d_feat, d_u = dev(torch, feat), dev(torch, u)
    d_wi = dev(torch, wi0)
    d_sizes = torch.zeros(8, dtype=torch.int64, device="cuda")
    rng = np.random.default_rng(92)
    for rep in range(12):
        nb = int(rng.integers(1, 9))
        sizes = [int(x) for x in rng.integers(0, 7500, nb)]
        for j in rng.choice(nb, size=min(nb, 2), replace=False):
            if rng.random() < 0.5:
                sizes[j] = 0
        n = sum(sizes)
        ref = _alone(torch, f, feat[:n], u[:n], wi0[:n], sizes, None)
        d_sizes.zero_()
        d_sizes[:nb].copy_(torch.tensor(sizes, dtype=torch.int64, device="cuda"))
        d_wi.copy_(dev(torch, wi0))
        wf.run(d_sizes[:nb], d_feat, d_u, d_wi)
        torch.cuda.synchronize()
        got = host(d_wi)
        assert np.array_equal(got[:n], ref), "replay %d sizes %s" % (rep, sizes)
        assert np.array_equal(got[n:], wi0[n:])
    wf.close()
    f.close()


def test_wavefront_autotune(torch_cuda):
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False))
    sizes = _bounces(30000, 6)
    cap = sum(sizes)
    feat, u = dev(torch, df_features(cap, 84)), dev(torch, uniforms(cap, 85))
    d_sizes = torch.tensor(sizes, dtype=torch.int64, device="cuda")
    wf = G.Wavefront(f, 6, cap)
    wi = dev(torch, wi_init(cap))
    best = wf.autotune(d_sizes, feat, u, wi)
    assert best["ms"] > 0 and best["blocks_per_cu"] in (2, 4, 6, 8, 12) and best["tiles_per_fetch"] in (1, 2, 4)
    wi2 = dev(torch, wi_init(cap))
    wf.run(d_sizes, feat, u, wi2)
    w3 = dev(torch, wi_init(cap))
    f.query(feat, u, w3)
    torch.cuda.synchronize()
    assert np.array_equal(host(wi2), host(w3))
    wf.close()
    f.close()


def _requests(torch, n_req, seed, nrc=False, lo=1024, hi=40000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_req)).astype(np.int64)
    reqs = []
    for i, n in enumerate(sizes):
        feat = df_features(int(n), seed * 1000 + i, nrc=nrc)
        u = uniforms(int(n), seed * 1000 + 500 + i)
        reqs.append((i, feat, u))
    return reqs


@pytest.mark.parametrize("n_streams,budget_rows,prepass_min", [(2, 1 << 22, 1 << 30), (1, 50000, 1 << 30),
                                                                (3, 120000, 1 << 30), (2, 1 << 22, 1024),
                                                                (3, 120000, 1024)])
def test_server_alone_equals_mix(torch_cuda, n_streams, budget_rows, prepass_min):
    """Every request equals the same request queried alone (alone: in-kernel
    grid gathers with prepass_min huge, the paired pre-pass with prepass_min
    1024; the server's rounds always run the pre-pass over their rows)."""
    torch = torch_cuda
    cfg = G.hash_config(with_nrc=True)
    cfg["dfg"]["grid_prepass_min_rows"] = prepass_min
    f = G.GuidingField(cfg)
    reqs = _requests(torch, 24, 90 + n_streams, nrc=True)
    alone = {}
    for i, feat, u in reqs:
        w = dev(torch, wi_init(len(feat)))
        rgb = torch.zeros((len(feat), 3), dtype=torch.float32, device="cuda")
        f.query(dev(torch, feat), dev(torch, u), w, nrc_rgb=rgb)
        alone[i] = (w, rgb)
    torch.cuda.synchronize()
    srv = G.Server(f, budget_bytes=budget_rows * 80, max_batch_requests=8, n_streams=n_streams, train_every=0)
    mixed = {}
    for i, feat, u in reqs:
        w = dev(torch, wi_init(len(feat)))
        rgb = torch.zeros((len(feat), 3), dtype=torch.float32, device="cuda")
        mixed[i] = (w, rgb)
        srv.submit(i, dev(torch, feat), dev(torch, u), w, nrc_rgb=rgb)
        srv.progress()
    srv.progress(wait=True)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["submitted"] == len(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    ids, lat, ver = srv.drain()
    assert sorted(ids.tolist()) == list(range(len(reqs))) and np.all(lat > 0)
    for i in alone:
        assert np.array_equal(host(alone[i][0]), host(mixed[i][0])), "request %d differs alone vs in the mix" % i
        assert np.array_equal(host(alone[i][1]), host(mixed[i][1]))
    srv.close()
    f.close()


def test_server_ring_wraps_and_budget_holds(torch_cuda):
    """5,000 small requests (more than the 4,096-entry request ring) through
    device-side admission under a 3,000-row budget: every request completes,
    each row equals one dfg_query over all rows (requests are slices of one
    buffer), and the rows in flight after any admission stay within the budget
    except for a single request larger than it, which runs alone."""
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False, log2_hashmap_size=14))
    rng = np.random.default_rng(21)
    sizes = rng.integers(1, 200, 5000)
    sizes[1234] = 4000                       # above the budget: admitted alone
    total = int(sizes.sum())
    feat, u = df_features(total, 22), uniforms(total, 23)
    d_feat, d_u = dev(torch, feat), dev(torch, u)
    ref = dev(torch, wi_init(total))
    f.query(d_feat, d_u, ref)
    budget_rows = 3000
    srv = G.Server(f, budget_bytes=budget_rows * 80, max_batch_requests=256, n_streams=2, train_every=0,
                   round_rows=1 << 16)
    out = dev(torch, wi_init(total))
    row = 0
    for i, n in enumerate(sizes.tolist()):
        srv.submit(i, d_feat[row:row + n], d_u[row:row + n], out[row:row + n])
        row += n
        if i % 500 == 0:
            srv.progress()
    srv.progress(wait=True)
    st = srv.stats()
    assert st["completed"] == len(sizes) and st["rows"] == total
    assert 0 < st["peak_rows"] <= max(budget_rows, int(sizes.max())), st["peak_rows"]
    ids, lat, _ = srv.drain(max_records=len(sizes))
    assert sorted(ids.tolist()) == list(range(len(sizes)))
    torch.cuda.synchronize()
    assert np.array_equal(host(out), host(ref))
    # a request above round_rows is refused at submission
    big = 1 << 16
    with pytest.raises(G.DfgError):
        srv.submit(99999, dev(torch, df_features(big + 1, 24)), dev(torch, uniforms(big + 1, 25)),
                   dev(torch, wi_init(big + 1)))
    srv.close()
    f.close()


@pytest.mark.parametrize("n_streams,budget_rows,max_req", [(1, 20000, 64), (2, 3000, 256), (4, 5000, 16)])
def test_server_many_requests_on_several_streams(torch_cuda, n_streams, budget_rows, max_req):
    """20,000 requests of 1..300 rows (several passes through the 4,096-entry
    ring) on 1, 2 or 4 launch streams, with rounds overlapping on the device
    and admission racing for the shared head: every request completes once,
    the rows in flight never exceed the budget, and every row equals one
    dfg_query over all rows."""
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False, log2_hashmap_size=14))
    rng = np.random.default_rng(31 + n_streams)
    sizes = rng.integers(1, 300, 20000)
    total = int(sizes.sum())
    d_feat, d_u = dev(torch, df_features(total, 32)), dev(torch, uniforms(total, 33))
    ref = dev(torch, wi_init(total))
    f.query(d_feat, d_u, ref)
    srv = G.Server(f, budget_bytes=budget_rows * 80, max_batch_requests=max_req, n_streams=n_streams, train_every=0,
                   round_rows=1 << 16)
    out = dev(torch, wi_init(total))
    row = 0
    for i, n in enumerate(sizes.tolist()):
        srv.submit(i, d_feat[row:row + n], d_u[row:row + n], out[row:row + n])
        row += n
        if rng.random() < 0.01:
            srv.progress()
    srv.progress(wait=True)
    st = srv.stats()
    assert st["completed"] == len(sizes) and st["rows"] == total
    assert 0 < st["peak_rows"] <= budget_rows, st["peak_rows"]
    ids, _, _ = srv.drain(max_records=len(sizes))
    assert sorted(ids.tolist()) == list(range(len(sizes)))
    torch.cuda.synchronize()
    assert np.array_equal(host(out), host(ref))
    srv.close()
    f.close()


def test_server_training_interleave_is_ordered(torch_cuda):
    """Requests launched before the first training step see the initial weights;
    requests launched after the last one see the final weights: no torn reads."""
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False))
    reqs = _requests(torch, 30, 7)
    before = {}
    for i, feat, u in reqs:
        w = dev(torch, wi_init(len(feat)))
        f.query(dev(torch, feat), dev(torch, u), w)
        before[i] = w
    torch.cuda.synchronize()
    B = 65536
    tf, tt = dev(torch, df_features(B, 8)), dev(torch, train_targets(B, 9))
    srv = G.Server(f, budget_bytes=1 << 30, max_batch_requests=2, n_streams=2, train_every=3)
    for _ in range(4):
        srv.submit_train(tf, tt)
    res = {}
    for i, feat, u in reqs:
        w = dev(torch, wi_init(len(feat)))
        res[i] = (w, dev(torch, feat), dev(torch, u))
        srv.submit(i, res[i][1], res[i][2], w)
    srv.progress(wait=True)
    ids, lat, ver = srv.drain()
    st = srv.stats()
    assert st["train_steps"] >= 1
    final = int(ver.max())
    assert final == st["train_steps"]
    for i, v in zip(ids.tolist(), ver.tolist()):
        if v == 0:
            assert np.array_equal(host(res[i][0]), host(before[i]))
        elif v == final:
            w = dev(torch, wi_init(res[i][0].shape[0]))
            f.query(res[i][1], res[i][2], w)
            torch.cuda.synchronize()
            assert np.array_equal(host(res[i][0]), host(w))
    srv.close()
    f.close()


# ----------------------------------------------------------------------------
# The two batched workloads at full size
# ----------------------------------------------------------------------------
def test_wavefront_full_frame_matches_per_bounce_and_fp64(torch_cuda):
    """The 6-bounce frame from 2,097,152 rows shrinking by 0.7 per bounce
    (6,168,076 queries), 64/64 bins, in one captured persistent launch with
    the grid pre-pass: bit-identical to dfg_query per bounce, and a 4,096-row
    slice (first rows of the first and last bounces) within the query's
    tolerance of the FP64 oracle."""
    import oracle as O
    torch = torch_cuda
    cfg = G.hash_config(n_bins_theta=64, n_bins_phi=64, with_nrc=False)
    f = G.GuidingField(cfg)
    sizes = _bounces(2097152, 6)
    assert sizes[-1] == 352467 and sum(sizes) == 6168076
    cap = sum(sizes)
    feat, u = df_features(cap, 180), uniforms(cap, 181)
    wi0 = wi_init(cap)
    d_feat, d_u = dev(torch, feat), dev(torch, u)
    ref = torch.empty((cap, 4), dtype=torch.float32, device="cuda")
    ref.copy_(dev(torch, wi0))
    row = 0
    for n in sizes:
        f.query(d_feat[row:row + n], d_u[row:row + n], ref[row:row + n])
        row += n
    wf = G.Wavefront(f, 6, cap)
    wf.set_tuning(8, 2, 1)
    d_sizes = torch.tensor(sizes, dtype=torch.int64, device="cuda")
    d_wi = dev(torch, wi0)
    wf.run(d_sizes, d_feat, d_u, d_wi)
    torch.cuda.synchronize()
    g, r = host(d_wi), host(ref)
    assert np.array_equal(g, r), "frame differs from per-bounce queries in %d rows" % (g != r).any(1).sum()
    th, ph, _ = O.config_layouts(cfg)
    rows = np.concatenate([np.arange(2048), cap - sizes[-1] + np.arange(2048)])
    o = O.query_f64(th, ph, f.get_params(0, G.DFG_PARAMS_INFERENCE), f.get_params(1, G.DFG_PARAMS_INFERENCE),
                    feat[rows], 9, u[rows], wi0[rows], 1)
    rel = np.abs(g[rows, 2] * g[rows, 3] - o[:, 2].astype(np.float64) * o[:, 3]) / np.abs(o[:, 2] * o[:, 3])
    assert np.median(rel) <= 1e-3 and np.percentile(rel, 99) <= 3e-2, (np.median(rel), np.percentile(rel, 99))
    wf.close()
    f.close()


def test_server_full_size_mix_alone_equals_mix(torch_cuda):
    """The served workload: 24 requests of log-uniform size in [1,024,
    1,048,576] rows, a 262,144-sample DF + NRC training step after every 8,
    a 16 GiB admission budget, two streams.  Every request equals the same
    request queried alone on a second field (same seed) that has taken the
    same training steps as the weights version the server reports for it;
    the pre-training requests also match the FP64 oracle on a slice."""
    import oracle as O
    torch = torch_cuda
    cfg = G.hash_config(with_nrc=True)
    fa, fb = G.GuidingField(cfg), G.GuidingField(cfg)
    rng = np.random.default_rng(4)
    n_req = 24
    sizes = np.exp(rng.uniform(np.log(1024), np.log(1 << 20), n_req)).astype(np.int64)
    sizes[3] = 1 << 20            # the largest size of the range is in the mix
    reqs = [(dev(torch, df_features(int(n), 2000 + i)), dev(torch, uniforms(int(n), 3000 + i)))
            for i, n in enumerate(sizes)]
    B = 262144
    tf = dev(torch, df_features(B, 8, nrc=True))
    tt = dev(torch, train_targets(B, 9))
    trgb = dev(torch, np.exp(np.random.default_rng(10).normal(0.0, 1.0, (B, 3))).astype(np.float32))
    srv = G.Server(fa, budget_bytes=16 << 30, max_batch_requests=64, n_streams=2, train_every=8)
    for _ in range(n_req // 8 + 1):
        srv.submit_train(tf, tt, nrc_feat=tf, nrc_target=trgb)
    outs = []
    for i, (fe, uu) in enumerate(reqs):
        w = dev(torch, wi_init(fe.shape[0]))
        outs.append(w)
        srv.submit(i, fe, uu, w)
        srv.progress()
    srv.progress(wait=True)
    ids, lat, ver = srv.drain()
    st = srv.stats()
    assert sorted(ids.tolist()) == list(range(n_req)) and st["train_steps"] >= 2
    version = dict(zip(ids.tolist(), ver.tolist()))
    final = max(version.values())
    assert final >= 1
    checked = 0
    for v in range(final + 1):
        for i in sorted(k for k in version if version[k] == v):
            fe, uu = reqs[i]
            w = dev(torch, wi_init(fe.shape[0]))
            fb.query(fe, uu, w)
            torch.cuda.synchronize()
            assert np.array_equal(host(outs[i]), host(w)), "request %d (version %d) differs alone vs mixed" % (i, v)
            checked += 1
            if v == 0 and i == min(k for k in version if version[k] == 0):
                th, ph, _ = O.config_layouts(cfg)
                k = min(4096, fe.shape[0])
                o = O.query_f64(th, ph, fb.get_params(0, G.DFG_PARAMS_INFERENCE),
                                fb.get_params(1, G.DFG_PARAMS_INFERENCE), host(fe[:k]), 9, host(uu[:k]),
                                wi_init(k), 1)
                g = host(outs[i])[:k]
                rel = np.abs(g[:, 2] * g[:, 3] - o[:, 2].astype(np.float64) * o[:, 3]) / np.abs(o[:, 2] * o[:, 3])
                assert np.median(rel) <= 1e-3, np.median(rel)
        if v < final:
            fb.train_step(tf, tt)
            fb.nrc_train_step(tf, trgb)
    assert checked == n_req
    srv.close()
    fa.close()
    fb.close()
