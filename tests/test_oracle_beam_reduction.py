"""The oracle's per-cell beam reduction (Grid::reduceBeams / Grid::intersect with
mMaxBeamsInCell > 0, SmallUPBP/src/Beams/Grid.hxx:318-399,440-631).

Pins: a bound above every cell's occupancy changes nothing; n identical beams crossing one
cell give the exact unreduced value under the fixed-count modes (max kept, each weighted
n / max); the reduced estimate averages to the unreduced one over seeds (unbiasedness); the
per-hit formulation (the cell of the query point, accel 0) agrees with the reference's
per-cell DDA (accel 1) apart from hits on cell boundaries."""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes

TYPES = [scenes.REDUCE_PRESAMPLE, scenes.REDUCE_OFFSET, scenes.REDUCE_RESAMPLE_FIXED, scenes.REDUCE_RESAMPLE]


def _oracle(w, res, max_in_cell=0, rtype=0, seed=1234):
    obm = po.OracleBeams(w.beams, grid_res=res)
    if max_in_cell:
        obm.set_reduction(max_in_cell, rtype, seed)
    return obm


@pytest.mark.parametrize("rtype", TYPES)
def test_bound_above_occupancy_changes_nothing(rtype):
    w = scenes.beams_stress(n_queries=400, n_beams=2000, seed=71)
    plain = _oracle(w, 16)
    red = _oracle(w, 16, 1 << 30, rtype)
    for accel in (0, 1):
        a, sa, _ = po.beams_eval(plain, w.media, w.queries, accel=accel, signatures=True)
        b, sb, _ = po.beams_eval(red, w.media, w.queries, accel=accel, signatures=True)
        assert np.array_equal(sa, sb) and np.array_equal(a, b)


@pytest.mark.parametrize("rtype", TYPES[:3])
def test_identical_beams_in_one_cell_known_answer(rtype):
    """n copies of one beam and a query crossing it: the fixed-count modes keep exactly max
    hits (counted with multiplicity) of weight n / max each, so the estimate equals the
    unreduced n C."""
    n, mx = 12, 5
    media = scenes.make_media([((0.1, 0.1, 0.1), (0.2, 0.2, 0.2), 0.0, 1.0)])
    bm = scenes.Beams(np.tile([[0.2, 0.5, 0.5]], (n, 1)).astype(np.float32),
                      np.tile([[1.0, 0.0, 0.0]], (n, 1)).astype(np.float32), np.full(n, 0.6, np.float32),
                      np.ones((n, 3), np.float32), np.full(n, scenes.SHORT_BEAM, np.uint32), np.zeros(n, np.int32),
                      0.05)
    d = np.array([[0.3, 0.0, 1.0]]) / np.hypot(0.3, 1.0)
    q = scenes.Queries(np.array([[0.35, 0.503, 0.0]], np.float32), d.astype(np.float32),
                       np.array([0, 1], np.int64),
                       scenes.segment_values(media, np.array([0], np.int32), np.array([0.0]), np.array([1.0])))
    plain = po.OracleBeams(bm, grid_res=8)
    a, sa, _ = po.beams_eval(plain, media, q, accel=0, signatures=True)
    assert sa[0, 0] == n and a[0, 0] > 0
    for seed in (1, 2, 3, 99):
        red = po.OracleBeams(bm, grid_res=8).set_reduction(mx, rtype, seed)
        for accel in (0, 1):
            b, sb, _ = po.beams_eval(red, media, q, accel=accel, signatures=True)
            assert sb[0, 0] == mx
            assert b[0] == pytest.approx(a[0], rel=2e-6)
    assert red.selection_pdf([0.5, 0.5, 0.5]) == pytest.approx(mx / n)


@pytest.mark.parametrize("rtype", TYPES)
def test_reduction_is_unbiased_over_seeds(rtype):
    w = scenes.beams_stress(n_queries=300, n_beams=3000, seed=72)
    plain = _oracle(w, 8)
    a, sa, _ = po.beams_eval(plain, w.media, w.queries, accel=0, signatures=True)
    total = a.astype(np.float64).sum()
    seeds = 160
    acc = 0.0
    kept = 0
    for s in range(seeds):
        red = _oracle(w, 8, 3, rtype, seed=1000 + s)
        b, sb, _ = po.beams_eval(red, w.media, w.queries, accel=0, signatures=True)
        acc += b.astype(np.float64).sum()
        kept += int(sb[:, 0].sum())
    mean = acc / seeds
    assert kept < 0.8 * seeds * sa[:, 0].sum()  # the reduction does drop hits
    assert abs(mean / total - 1.0) < 0.03, (mean, total)


@pytest.mark.parametrize("rtype", TYPES)
def test_per_hit_cells_agree_with_the_reference_dda(rtype):
    w = scenes.beams_stress(n_queries=1500, n_beams=6000, seed=33)
    obm = _oracle(w, 24, 2, rtype, seed=5)
    a, sa, ca = po.beams_eval(obm, w.media, w.queries, accel=0, signatures=True)
    b, sb, cb = po.beams_eval(obm, w.media, w.queries, accel=1, signatures=True)
    same = np.all(sa == sb, axis=1)
    assert (sa[:, 0] > 0).sum() > 200
    # hits on a cell boundary may fall to the neighbouring cell in one formulation
    assert same.mean() > 0.99
    assert np.allclose(a[same], b[same], rtol=1e-5, atol=0)


def test_reduction_needs_the_grid_and_a_valid_type():
    w = scenes.beams_stress(n_queries=10, n_beams=50, seed=73)
    with pytest.raises(RuntimeError):
        po.OracleBeams(w.beams, grid_res=0).set_reduction(2, 0, 1)
    with pytest.raises(RuntimeError):
        po.OracleBeams(w.beams, grid_res=8).set_reduction(2, 7, 1)
