"""dfg_query_noalloc (the allocation-free query) against dfg_query, through
the C ABI: bit-identical results on consecutive calls of shrinking size (the
second and third reuse the device scratch array the first filled), with and
without the radiance cache, on the paired hash grids and on the reference's
DenseGrid F = 4 configuration (per-net pre-pass); no device memory is taken
by the calls; sizes beyond the scratch are rejected before any launch."""
import numpy as np
import pytest

import dfguide as G
from gpu_util import df_features, uniforms, wi_init, dev, host

pytestmark = pytest.mark.gpu

MAX_ROWS_HASH = 65536  # two 16-level F = 2 grids in the 16.75 MiB scratch array


def _run(torch, f, fn, feat, u, guided, with_rgb, free_before=None):
    """One call of fn on device copies of the inputs; with free_before (a
    list), appends the device's free bytes just before and after the call
    (the inputs and outputs are allocated by torch beforehand)."""
    n = feat.shape[0]
    d_feat, d_u, d_g = dev(torch, feat), dev(torch, u), dev(torch, guided)
    d_wi = dev(torch, wi_init(n))
    d_rgb = torch.zeros((n, 3), dtype=torch.float32, device="cuda") if with_rgb else None
    torch.cuda.synchronize()
    if free_before is not None:
        free_before.append(torch.cuda.mem_get_info()[0])
    fn(d_feat, d_u, d_wi, guided=d_g, nrc_rgb=d_rgb)
    torch.cuda.synchronize()
    if free_before is not None:
        free_before.append(torch.cuda.mem_get_info()[0])
    return host(d_wi), (host(d_rgb) if with_rgb else None)


@pytest.mark.parametrize("cfgname", ["hash", "reference"])
def test_noalloc_query_equals_query(torch_cuda, cfgname):
    torch = torch_cuda
    cfg = G.hash_config(with_nrc=True) if cfgname == "hash" else G.reference_config(with_nrc=True)
    f = G.GuidingField(cfg)
    for k, n in enumerate([MAX_ROWS_HASH, 40000, 1000]):
        feat, u = df_features(n, 40 + k, nrc=True), uniforms(n, 50 + k)
        guided = (np.random.default_rng(60 + k).random(n) < 0.8).astype(np.uint8)
        with_rgb = k != 1
        free = []
        a_wi, a_rgb = _run(torch, f, f.query_noalloc, feat, u, guided, with_rgb, free)
        # the call takes no device memory (dfg_query's pre-pass buffers for
        # 65,536 rows would take 17 MB)
        assert free[1] >= free[0] - (1 << 20), "call %d took %d bytes" % (k, free[0] - free[1])
        b_wi, b_rgb = _run(torch, f, f.query, feat, u, guided, with_rgb)
        assert np.array_equal(a_wi, b_wi), "call %d (n=%d): wi differs in %d rows" % (
            k, n, (a_wi != b_wi).any(1).sum())
        if with_rgb:
            assert np.array_equal(a_rgb, b_rgb), "call %d: NRC radiance differs" % k
        assert np.all(np.isfinite(a_wi)) and np.all(a_wi[:, 2] * a_wi[:, 3] > 0)
    f.close()


def test_noalloc_query_rejects_rows_beyond_the_scratch(torch_cuda):
    torch = torch_cuda
    f = G.GuidingField(G.hash_config(with_nrc=False))
    n = MAX_ROWS_HASH + 64
    feat = torch.zeros((n, 9), dtype=torch.float32, device="cuda")
    u = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    wi = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(G.DfgError) as e:
        f.query_noalloc(feat, u, wi)
    assert e.value.code == 2 and "65536 rows" in str(e.value)
    # the context still works after the rejection
    m = 1000
    a_wi, _ = _run(torch, f, f.query_noalloc, df_features(m, 70), uniforms(m, 71), np.ones(m, np.uint8), False)
    b_wi, _ = _run(torch, f, f.query, df_features(m, 70), uniforms(m, 71), np.ones(m, np.uint8), False)
    assert np.array_equal(a_wi, b_wi)
    f.close()
