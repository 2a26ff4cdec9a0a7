"""upbp_eval_noalloc (the allocation-free evaluation) against upbp_eval through the C ABI.

Both run the same kernels on the same device-resident queries; the allocation-free call keeps
its counters, query keys, permutations and sort storage in the library's fixed per-device
region instead of the context's buffers. The outputs and hit counters must be identical:
two calls in a row, the second with a smaller batch (the region keeps the first call's
contents, so what must start at zero is cleared by the call itself). Batches the region
cannot hold and host-resident inputs are rejected with UPBP_EINVAL.
"""
import numpy as np
import pytest
import torch

from smallupbp_amd import _abi, native, scenes

from gpu_helpers import gpu_context, print_case

pytestmark = pytest.mark.gpu

MAX_SORTED = 1 << 21  # UPBP_NOALLOC_MAX_SORTED_QUERIES


def _both(ctx, bre, bm, qv):
    """upbp_eval with device outputs, then upbp_eval_noalloc into fresh device outputs."""
    s1, s2 = {}, {}
    ep, eb = native.eval_batch(ctx, bre, bm, qv, stats=s1)
    n = qv.count
    np_ = torch.full((n, 3), -7.0, device="cuda")
    nb = torch.full((n, 3), -7.0, device="cuda")
    native.eval_batch_noalloc(ctx, bre, bm, qv, np_, nb, stats=s2)
    torch.cuda.synchronize()
    return (ep, eb, s1), (np_, nb, s2)


def _check_same(a, b):
    (ep, eb, s1), (np_, nb, s2) = a, b
    assert torch.equal(ep, np_), "photon radiance differs between upbp_eval and upbp_eval_noalloc"
    assert torch.equal(eb, nb), "beam radiance differs between upbp_eval and upbp_eval_noalloc"
    for k in ("photon_hits", "beam_hits", "photon_candidates", "beam_candidates", "nodes_visited"):
        assert s1[k] == s2[k], (k, s1[k], s2[k])
    assert s1["photon_hits"] > 0 and s1["beam_hits"] > 0


def test_noalloc_equals_eval_twice_second_smaller():
    print_case("noalloc_vs_eval", queries="256x144 then 96x54", photons=100_000, beams=10_000)
    w = scenes.bre_beams(256, 144, 100_000, 10_000, seed=11)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    big = native.QueryBatchView.to_device(w.queries)
    _check_same(*_both(ctx, bre, bm, big))
    # second call: a smaller batch (every third query) that does not fill what the first left
    idx = np.arange(0, w.queries.count, 3)
    small = native.QueryBatchView.to_device(w.queries.subset(idx))
    a, b = _both(ctx, bre, bm, small)
    _check_same(a, b)
    # and the unsorted order of the same small batch
    small_as_given = native.QueryBatchView.to_device(w.queries.subset(idx), order=_abi.ORDER_AS_GIVEN)
    _check_same(*_both(ctx, bre, bm, small_as_given))


def test_noalloc_photons_only_then_beams_only():
    print_case("noalloc_single_estimator")
    w = scenes.bre_beams(64, 36, 20_000, 2_000, seed=12)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries)
    ref_p, _ = native.eval_batch(ctx, bre, None, qv)
    _, ref_b = native.eval_batch(ctx, None, bm, qv)
    out = torch.zeros((qv.count, 3), device="cuda")
    native.eval_batch_noalloc(ctx, bre, None, qv, out, None)
    assert torch.equal(out, ref_p)
    native.eval_batch_noalloc(ctx, None, bm, qv, None, out)
    assert torch.equal(out, ref_b)


def test_noalloc_rejects_what_it_cannot_hold():
    print_case("noalloc_rejects")
    w = scenes.bre_beams(32, 18, 5_000, 500, seed=13)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    out = torch.zeros((MAX_SORTED + 1, 3), device="cuda")
    # host-resident queries need staging copies: rejected
    with pytest.raises(native.UpbpError) as e:
        native.eval_batch_noalloc(ctx, bre, None, native.QueryBatchView.host(w.queries), out, None)
    assert e.value.status == _abi.UPBP_EINVAL
    # a sorted batch above the region's limit: rejected before any device work
    n = MAX_SORTED + 1
    o = torch.zeros((n, 3), device="cuda")
    d = torch.zeros((n, 3), device="cuda")
    d[:, 2] = 1.0
    off = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    seg = torch.zeros((1, 8), device="cuda")
    big = native.QueryBatchView(o, d, off, seg)
    with pytest.raises(native.UpbpError) as e:
        native.eval_batch_noalloc(ctx, bre, None, big, out, None)
    assert e.value.status == _abi.UPBP_EINVAL
    assert "UPBP_NOALLOC_MAX_SORTED_QUERIES" in e.value.message
    # the same batch unsorted is accepted (segment-free queries: zero radiance)
    big_as_given = native.QueryBatchView(o, d, off, seg, order=_abi.ORDER_AS_GIVEN)
    out.fill_(-1.0)
    native.eval_batch_noalloc(ctx, bre, None, big_as_given, out, None)
    torch.cuda.synchronize()
    assert bool((out == 0).all())
