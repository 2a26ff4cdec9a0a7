"""GPU parity of the per-cell beam reduction (max_beams_in_cell > 0) against the CPU oracle.

The GPU registers the beams in the reference's uniform grid and decides each hit by the cell
of its query point; the oracle does the same (accel 0) with its own grid build, and the
reference's per-cell DDA formulation (accel 1) is checked against that on the CPU. Bar: hit
multisets identical (count and hash weighted by multiplicity), values within 1e-4 (relative,
floored at the median query value for radius-grazing hits), the cells' selection pdfs
identical."""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import _abi, native, scenes

from gpu_helpers import REL_TOL, gpu_context, print_case, rel_err

pytestmark = pytest.mark.gpu

TYPES = [scenes.REDUCE_PRESAMPLE, scenes.REDUCE_OFFSET, scenes.REDUCE_RESAMPLE_FIXED, scenes.REDUCE_RESAMPLE]


def _gpu(w, mx, rtype, res, seed, ctx=None):
    ctx = ctx or gpu_context(w.media)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, max_beams_in_cell=mx, reduction_type=rtype, grid_resolution=res, reduction_seed=seed)
    qv = native.QueryBatchView.to_device(w.queries)
    st = {}
    out = bm.evalBeamBeamEstimate(qv, stats=st).cpu().numpy()
    _, sig = native.hit_signatures(ctx, None, bm, qv)
    return ctx, bm, out, sig, st


def _oracle(w, mx, rtype, res, seed):
    obm = po.OracleBeams(w.beams, grid_res=res)
    if mx:
        obm.set_reduction(mx, rtype, seed)
    return obm


@pytest.mark.parametrize("rtype", TYPES)
def test_reduced_hits_and_values_match_oracle(rtype):
    print_case("beam_reduction", rtype=rtype)
    w = scenes.beams_stress(n_queries=3000, n_beams=12_000, seed=81)
    ctx, bm, out, sig, st = _gpu(w, 2, rtype, 24, 77)
    obm = _oracle(w, 2, rtype, 24, 77)
    ref, rsig, cnt = po.beams_eval(obm, w.media, w.queries, accel=0, signatures=True)
    plain, psig, _ = po.beams_eval(_oracle(w, 0, 0, 24, 0), w.media, w.queries, accel=0, signatures=True)
    diff = np.any(sig != rsig, axis=1)
    print("hits kept %d of %d; queries differing %d" % (rsig[:, 0].sum(), psig[:, 0].sum(), diff.sum()))
    assert rsig[:, 0].sum() < 0.6 * psig[:, 0].sum()
    assert not diff.any(), "reduced beam hit sets differ"
    assert st["beam_hits"] == int(cnt[0])
    # A hit grazing the radius (1 - d^2/r^2 near 0) carries the cancellation of the kernel: an
    # ulp of distance is a large relative change of a tiny value. With few kept hits such a
    # hit can be a query's whole value, so the bar floors the denominator at the median query
    # value (an absolute 1e-4 of a typical query) and counts the strictly relative outliers.
    mag = np.abs(ref).max(axis=1)
    floor = float(np.median(mag[mag > 0]))
    strict = rel_err(out, ref)
    print("queries above %.0e relative: %d" % (REL_TOL, (strict > REL_TOL).sum()))
    assert rel_err(out, ref, floor=floor).max() <= REL_TOL
    assert (strict > REL_TOL).mean() < 1e-3


def test_selection_pdf_matches_oracle_grid():
    print_case("beam_reduction_pdf")
    w = scenes.beams_stress(n_queries=10, n_beams=12_000, seed=82)
    ctx = gpu_context(w.media)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, max_beams_in_cell=3, reduction_type=scenes.REDUCE_RESAMPLE, grid_resolution=32)
    obm = _oracle(w, 3, scenes.REDUCE_RESAMPLE, 32, 1234)
    rng = np.random.default_rng(5)
    pts = np.concatenate([rng.uniform(-0.2, 1.2, (300, 3)),
                          w.beams.origin[:300] + w.beams.direction[:300] * (0.5 * w.beams.length[:300, None])])
    g = np.array([bm.getBeamSelectionPdf(p) for p in pts.astype(np.float32)])
    o = np.array([obm.selection_pdf(p) for p in pts.astype(np.float32)])
    print("cells reduced among the probes: %d" % (o < 1).sum())
    assert (o < 1).sum() > 50
    assert np.array_equal(g, o)
    # the batched entry point, from host and from device memory
    assert np.array_equal(bm.getBeamSelectionPdfs(pts), o)
    import torch

    dev = bm.getBeamSelectionPdfs(torch.from_numpy(pts.astype(np.float32)).cuda())
    assert np.array_equal(dev.cpu().numpy(), o)
    plain = native.PhotonBeamsEvaluator(ctx)
    plain.build(w.beams)
    assert np.all(plain.getBeamSelectionPdfs(pts) == 1.0)


def test_bound_above_occupancy_equals_unreduced():
    print_case("beam_reduction_bound")
    w = scenes.beams_stress(n_queries=2000, n_beams=8_000, seed=83)
    ctx, _, a, sa, _ = _gpu(w, 0, 0, 0, 0)
    _, _, b, sb, _ = _gpu(w, 1 << 30, scenes.REDUCE_OFFSET, 16, 3, ctx=ctx)
    assert np.array_equal(sa, sb) and np.array_equal(a, b)


def test_reduction_rejected_where_unsupported():
    """A reduced beam grid in a batched plan is rejected (UPBP_EUNSUPPORTED); the adjoint and
    the UPBP weights accept it (their parity tests are in test_gpu_adjoint.py and
    test_gpu_upbp.py); an unknown reduction type is rejected at build."""
    import torch

    print_case("beam_reduction_rejects")
    w = scenes.beams_stress(n_queries=64, n_beams=500, seed=84)
    ctx = gpu_context(w.media)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, max_beams_in_cell=2, reduction_type=scenes.REDUCE_PRESAMPLE, grid_resolution=8)
    qv = native.QueryBatchView.to_device(w.queries)
    with pytest.raises(native.UpbpError) as e:
        native.BatchPlan(ctx, [(None, bm, qv, 0, None, torch.zeros((w.queries.count, 3), device="cuda"))])
    assert e.value.status == _abi.UPBP_EUNSUPPORTED
    g = np.zeros((w.queries.count, 3), np.float32)
    native.adjoint(ctx, None, bm, qv, grad_out_beams=g, grad_scattering=np.zeros((len(w.media), 3), np.float32),
                   grad_absorption=np.zeros((len(w.media), 3), np.float32))
    with pytest.raises(native.UpbpError):
        bad = native.PhotonBeamsEvaluator(ctx)
        bad.build(w.beams, max_beams_in_cell=2, reduction_type=9)
