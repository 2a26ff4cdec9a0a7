"""GPU parity of the per-query grid statistics (upbp_beams_grid_stats: GridStats,
SmallUPBP/src/Beams/GridStats.hxx:37-54, of Grid::intersect's walk, Grid.hxx:318-399, over a
query's scattering segments, PhBeams.cxx:177-230) against the CPU oracle.

Both walk the reference's grid with its FP32 arithmetic, so the counts are integers that must
be identical: on the GPU-built grid of a reduction (max_beams_in_cell > 0), on the grid built
for the call when the structure has none, for LONG and SHORT queries, host and device query
batches, and on a sample of the 1920x1080 workload's queries."""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import native, scenes
from gpu_helpers import gpu_context, print_case

pytestmark = pytest.mark.gpu


def _oracle_stats(w, res, mx=0, queries=None):
    obm = po.OracleBeams(w.beams, grid_res=res)
    if mx:
        obm.set_reduction(mx, scenes.REDUCE_PRESAMPLE, 11)
    return po.beam_grid_stats(obm, w.media, queries or w.queries)


def _gpu_stats(w, res, mx=0, queries=None, host=False):
    ctx = gpu_context(w.media)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, max_beams_in_cell=mx, reduction_type=scenes.REDUCE_PRESAMPLE, grid_resolution=res,
             reduction_seed=11)
    q = queries or w.queries
    qv = native.QueryBatchView.host(q) if host else native.QueryBatchView.to_device(q)
    return bm.grid_stats(qv)


def _compare(got, want, case):
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, "%s: %d of %d queries differ, first %s: gpu %s oracle %s" % (
        case, len(bad), len(got), bad[:3], got[bad[:3]].tolist(), want[bad[:3]].tolist())


@pytest.mark.parametrize("res,mx", [(16, 0), (16, 3), (40, 0), (40, 2)])
def test_box_sphere_scene(res, mx):
    """Camera rays through the box and the nested sphere: up to three scattering segments."""
    print_case("grid_stats box_sphere", res=res, max_in_cell=mx)
    w = scenes.bre_beams(64, 36, 100, 6000, seed=91)
    want = _oracle_stats(w, res, mx)
    got = _gpu_stats(w, res, mx)
    _compare(got, want, "res %d max %d" % (res, mx))
    assert want[:, 1].sum() > 0 and (mx == 0 or want[:, 2].sum() > 0)


@pytest.mark.parametrize("host", [False, True])
def test_short_queries_and_host_batches(host):
    print_case("grid_stats short", host=host)
    w = scenes.bre_beams(48, 27, 100, 4000, seed=92)
    q = dataclasses.replace(w.queries, beam_type=scenes.SHORT_BEAM)
    want = _oracle_stats(w, 24, 4, queries=q)
    got = _gpu_stats(w, 24, 4, queries=q, host=host)
    _compare(got, want, "SHORT host=%s" % host)


def test_stress_scene_default_resolution():
    """The reference's default 256-cell grid (grid_resolution 0) over 20k mixed SHORT / LONG beams."""
    print_case("grid_stats stress", res=0)
    w = scenes.beams_stress(n_queries=3000, n_beams=20000, seed=93)
    want = _oracle_stats(w, 256)
    got = _gpu_stats(w, 0)
    _compare(got, want, "stress")
    assert want[:, 1].max() > 100


def test_1080p_workload_sample():
    """200k beams of the 1920x1080 workload; 20k random queries of its 2.07M."""
    print_case("grid_stats 1080p sample")
    w = scenes.bre_beams(1920, 1080, 1000, 200_000)
    idx = np.sort(np.random.default_rng(94).choice(w.queries.count, 20_000, replace=False))
    q = w.queries.subset(idx)
    want = _oracle_stats(w, 256, queries=q)
    got = _gpu_stats(w, 0, queries=q)
    _compare(got, want, "1080p sample")


def test_no_beams_and_bad_medium():
    w = scenes.bre_beams(16, 9, 100, 500, seed=95)
    ctx = gpu_context(w.media)
    empty = dataclasses.replace(w.beams, origin=w.beams.origin[:0], direction=w.beams.direction[:0],
                                length=w.beams.length[:0], throughput=w.beams.throughput[:0],
                                flags=w.beams.flags[:0], medium_id=w.beams.medium_id[:0])
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(empty)
    assert not bm.grid_stats(native.QueryBatchView.to_device(w.queries)).any()
    # a segment medium outside the table is an error, not a read past the media
    seg = w.queries.segments.copy()
    seg["medium_id"][0] = 57
    bad = dataclasses.replace(w.queries, segments=seg)
    bm2 = native.PhotonBeamsEvaluator(ctx)
    bm2.build(w.beams)
    with pytest.raises(native.UpbpError):
        bm2.grid_stats(native.QueryBatchView.to_device(bad))
