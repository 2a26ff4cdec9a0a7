"""GPU parity of the UPBP (MIS-weighted) estimators against the CPU oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import native, scenes

from gpu_helpers import REL_TOL, gpu_context, print_case, rel_err

pytestmark = pytest.mark.gpu

ALL = scenes.PB2D | scenes.BB1D | scenes.BPT | scenes.PP3D | scenes.SURF


def _gpu(w, inm, mi, device=True):
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons, in_medium=inm)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries) if device else native.QueryBatchView.host(w.queries)
    op, ob = native.eval_upbp(ctx, bre, bm, qv, mi)
    if hasattr(op, "cpu"):
        op, ob = op.cpu().numpy(), ob.cpu().numpy()
    return op, ob


def _oracle(w, inm, mi):
    oph = po.OraclePhotons(w.photons, in_medium=inm, build_bvh=False)
    obm = po.OracleBeams(w.beams, grid_res=0)
    a, _ = po.upbp_eval(oph, None, w.media, w.queries, mi, 0)
    b, _ = po.upbp_eval(None, obm, w.media, w.queries, mi, 1)
    return a, b


CASES = [
    ("all_long_short", ALL, scenes.LONG_BEAM, scenes.SHORT_BEAM, False),
    ("all_long_long", ALL, scenes.LONG_BEAM, scenes.LONG_BEAM, False),
    ("all_short_short", ALL, scenes.SHORT_BEAM, scenes.SHORT_BEAM, False),
    ("no_sine", ALL | scenes.NO_SINE_IN_WEIGHTS, scenes.LONG_BEAM, scenes.SHORT_BEAM, False),
    ("previous", ALL | scenes.PREVIOUS, scenes.LONG_BEAM, scenes.SHORT_BEAM, False),
    ("bb1d_previous", ALL | scenes.BB1D_PREVIOUS, scenes.LONG_BEAM, scenes.SHORT_BEAM, False),
    ("lite_segments", ALL, scenes.LONG_BEAM, scenes.SHORT_BEAM, True),
]


@pytest.mark.parametrize("name,tech,qbt,pbt,lite", CASES)
def test_upbp_matches_oracle(name, tech, qbt, pbt, lite):
    print_case("upbp", name=name)
    w, inm, mi = scenes.upbp_scene(40, 24, 3000, 800, seed=61, techniques=tech, query_beam_type=qbt,
                                   photon_beam_type=pbt)
    w.queries.lite = lite
    gp, gb = _gpu(w, inm, mi)
    rp, rb = _oracle(w, inm, mi)
    assert (rp.max(1) > 0).sum() > 100 and (rb.max(1) > 0).sum() > 50
    ep, eb = rel_err(gp, rp), rel_err(gb, rb)
    print("max rel err photons %.2e beams %.2e" % (ep.max(), eb.max()))
    assert ep.max() <= REL_TOL and eb.max() <= REL_TOL


def test_upbp_host_inputs_and_bad_indices():
    print_case("upbp_host_errors")
    w, inm, mi = scenes.upbp_scene(24, 14, 1500, 400, seed=62)
    gp, gb = _gpu(w, inm, mi, device=False)
    rp, rb = _oracle(w, inm, mi)
    assert rel_err(gp, rp).max() <= REL_TOL and rel_err(gb, rb).max() <= REL_TOL
    bad = scenes.MisInputs(mi.queries.copy(), mi.camera, mi.light_vertices, mi.path_ends, mi.beams, mi.params)
    bad.queries["camera_offset"][3] = len(mi.camera)  # past the camera array
    with pytest.raises(native.UpbpError):
        _gpu(w, inm, bad)


def test_upbp_long_paths_zero_pdfs_and_malformed_lengths():
    """Light sub-paths up to 8 segments with zero reverse pdfs (the walks stop there) and a few
    paths whose stored path lengths do not count up by one (the walks then read the vertices
    below by position, as the reference does)."""
    print_case("upbp_edge_paths")
    w, inm, mi = scenes.upbp_scene(40, 24, 3000, 800, seed=63, techniques=ALL, max_light_length=8)
    lv = mi.light_vertices.copy()
    rng = np.random.default_rng(5)
    m = lv["mis"]
    z = rng.uniform(size=len(lv))
    m["rev_pdf_a"][z < 0.08] = 0.0
    m["rev_pdf_a_without_bsdf"][(z >= 0.08) & (z < 0.16)] = 0.0
    starts = np.concatenate([[0], mi.path_ends[:-1]]).astype(np.int64)
    lens = mi.path_ends - starts
    bad = [p for p in rng.choice(len(starts) - 1, 200, replace=False) if lens[p] >= 4]
    for p in bad:  # a middle vertex claims one more segment (still inside the array)
        lv["path_length"][starts[p] + 2] += 1
    mi = scenes.MisInputs(mi.queries, mi.camera, lv, mi.path_ends, mi.beams, mi.params)
    gp, gb = _gpu(w, inm, mi)
    rp, rb = _oracle(w, inm, mi)
    assert len(bad) > 50 and (rp.max(1) > 0).sum() > 100 and (rb.max(1) > 0).sum() > 50
    ep, eb = rel_err(gp, rp), rel_err(gb, rb)
    print("max rel err photons %.2e beams %.2e" % (ep.max(), eb.max()))
    assert ep.max() <= REL_TOL and eb.max() <= REL_TOL


@pytest.mark.parametrize("pbt", [scenes.LONG_BEAM, scenes.SHORT_BEAM])
def test_upbp_long_camera_paths_bb1d_previous(pbt):
    """Camera sub-paths of up to 6 vertices (the per-query camera split past index 1) with
    BB1D_PREVIOUS, whose resets drop earlier BB1D terms."""
    print_case("upbp_long_camera_paths", pbt=pbt)
    w, inm, mi = scenes.upbp_scene(40, 24, 3000, 800, seed=64, techniques=ALL | scenes.BB1D_PREVIOUS,
                                   query_beam_type=scenes.LONG_BEAM, photon_beam_type=pbt, max_camera_length=6)
    gp, gb = _gpu(w, inm, mi)
    rp, rb = _oracle(w, inm, mi)
    assert (rp.max(1) > 0).sum() > 100 and (rb.max(1) > 0).sum() > 50
    ep, eb = rel_err(gp, rp), rel_err(gb, rb)
    print("max rel err photons %.2e beams %.2e" % (ep.max(), eb.max()))
    assert ep.max() <= REL_TOL and eb.max() <= REL_TOL


@pytest.mark.parametrize("rtype", [scenes.REDUCE_PRESAMPLE, scenes.REDUCE_OFFSET, scenes.REDUCE_RESAMPLE_FIXED,
                                   scenes.REDUCE_RESAMPLE])
def test_upbp_with_reduced_beam_cells_matches_oracle(rtype):
    """The photon-beam evaluator reduces its grid cells (max 2 beams tested per cell of an 8-cell
    grid): beam hits are kept as the cell of their query point decides and weighted with the
    cell's selection pdf in the contribution and the MIS weights; PB2D hits take the pdf of the
    cell holding their hit point as the camera vertex's BB1D selection pdf. Same values as the
    oracle (relative, floored at the median query for radius-grazing beam hits, as the plain
    reduction test), identical kept-hit counts."""
    print_case("upbp_reduced", rtype=rtype)
    w, inm, mi = scenes.upbp_scene(80, 48, 3000, 800, seed=65, techniques=ALL)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons, in_medium=inm)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, max_beams_in_cell=2, reduction_type=rtype, grid_resolution=8, reduction_seed=41)
    qv = native.QueryBatchView.to_device(w.queries)
    st = {}
    gp, gb = native.eval_upbp(ctx, bre, bm, qv, mi, stats=st)
    gp, gb = gp.cpu().numpy(), gb.cpu().numpy()
    oph = po.OraclePhotons(w.photons, in_medium=inm, build_bvh=False)
    obm = po.OracleBeams(w.beams, grid_res=8).set_reduction(2, rtype, 41)
    rp, hp = po.upbp_eval(oph, obm, w.media, w.queries, mi, 0)
    rb, hb = po.upbp_eval(None, obm, w.media, w.queries, mi, 1)
    plain_b, _ = po.upbp_eval(None, po.OracleBeams(w.beams, grid_res=0), w.media, w.queries, mi, 1)
    sel = np.array([obm.selection_pdf(p) for p in w.photons.position[inm.astype(bool)][:500]])
    print("photon hits %d / %d, beam tests %d / %d; reduced cells among 500 photon points: %d"
          % (st["photon_hits"], hp, st["beam_hits"], hb, (sel < 1).sum()))
    assert (sel < 1).sum() > 100
    assert st["beam_hits"] == hb and st["photon_hits"] == hp
    assert (rb.max(1) > 0).sum() > 50 and np.abs(rb - plain_b).max() > 0
    ep = rel_err(gp, rp)
    mag = np.abs(rb).max(axis=1)
    floor = float(np.median(mag[mag > 0]))
    eb = rel_err(gb, rb, floor=floor)
    print("max rel err photons %.2e beams %.2e (strict beams %.2e)" % (ep.max(), eb.max(), rel_err(gb, rb).max()))
    assert ep.max() <= REL_TOL and eb.max() <= REL_TOL
    assert (rel_err(gb, rb) > REL_TOL).mean() <= 1e-3
    # without out_beams the reduced structure only gives PB2D hits their selection pdf
    gp2, gb2 = native.eval_upbp(ctx, bre, bm, qv, mi, photons_only=True)
    assert gb2 is None and np.array_equal(gp2.cpu().numpy(), gp)
