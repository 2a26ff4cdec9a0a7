"""GPU parity of the KNN_RADIUS builds against the oracle (bit-exact radii).

Reference: EmbreeBre::build (SmallUPBP/src/Bre/Bre.cxx:562-602) and
PhotonBeamsEvaluator::build (SmallUPBP/src/Beams/PhBeams.cxx:77-117).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import native, scenes

from gpu_helpers import REL_TOL, gpu_context, print_case, rel_err

pytestmark = pytest.mark.gpu


def _clustered_photons(n, seed):
    # half uniform in the unit box, half in a dense ball: both sparse and crowded neighbourhoods
    rng = scenes.rng_for(seed)
    a = rng.random((n // 2, 3), dtype=np.float32)
    b = (0.5 + 0.05 * rng.standard_normal((n - n // 2, 3))).astype(np.float32)
    pos = np.concatenate([a, b]).astype(np.float32)
    inc = scenes._isotropic_dirs(rng, n).astype(np.float32)
    power = (rng.random((n, 3), dtype=np.float32) * 1e-4).astype(np.float32)
    return scenes.Photons(pos, inc, power, None, 0.0)


@pytest.mark.parametrize("knn", [2, 8, 10, 16, 17, 40, 150])
def test_photon_knn_radii_bit_exact(knn):
    print_case("photon_knn", knn=knn)
    w = scenes.plain_bre(16, 16, 100, 0.01, seed=5)
    ph = _clustered_photons(20_000, 11)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(ph, radius=0.5, radius_calculation=native._abi.KNN_RADIUS, knn=knn)
    got = bre.radii()
    ref = po.photon_knn_radius(ph.position, 0.5, knn)
    assert got.dtype == np.float32 and np.array_equal(got, ref)


def test_photon_knn_with_in_medium_filter_and_few_photons():
    print_case("photon_knn_filter")
    w = scenes.plain_bre(16, 16, 100, 0.01, seed=5)
    ph = _clustered_photons(6_000, 12)
    inm = (np.arange(ph.count) % 3 != 0).astype(np.uint8)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    n = bre.build(ph, radius=0.25, in_medium=inm, radius_calculation=native._abi.KNN_RADIUS, knn=12)
    assert n == int(inm.sum())
    assert np.array_equal(bre.radii(), po.photon_knn_radius(ph.position, 0.25, 12, in_medium=inm))
    # fewer kept photons than knn: the farthest one sets the radius
    few = np.zeros(ph.count, np.uint8)
    few[[3, 500, 4000, 5999]] = 1
    bre2 = native.EmbreeBre(ctx)
    bre2.build(ph, radius=1.0, in_medium=few, radius_calculation=native._abi.KNN_RADIUS, knn=9)
    assert np.array_equal(bre2.radii(), po.photon_knn_radius(ph.position, 1.0, 9, in_medium=few))
    # knn = 1 is the photon itself: radius 0 is rejected
    bre3 = native.EmbreeBre(ctx)
    with pytest.raises(native.UpbpError):
        bre3.build(ph, radius=1.0, radius_calculation=native._abi.KNN_RADIUS, knn=1)


@pytest.mark.parametrize("knn", [3, 12, 30])
def test_beam_knn_radii_bit_exact(knn):
    print_case("beam_knn", knn=knn)
    w = scenes.bre_beams(32, 18, 1000, 20_000, seed=13)
    ctx = gpu_context(w.media)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, radius=0.2, radius_calculation=native._abi.KNN_RADIUS, knn=knn)
    g0, g1 = bm.radii()
    r0, r1 = po.beam_knn_radius(w.beams, 0.2, knn)
    assert np.array_equal(g0, r0) and np.array_equal(g1, r1)


def test_knn_estimates_match_oracle_with_those_radii():
    print_case("knn_estimates")
    w = scenes.bre_beams(96, 54, 40_000, 8_000, seed=14)
    # KNN radii replace the scene's per-photon radii
    w.photons = scenes.Photons(w.photons.position, w.photons.inc_dir, w.photons.power, None, 0.0)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons, radius=0.3, radius_calculation=native._abi.KNN_RADIUS, knn=10)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams, radius=0.05, radius_calculation=native._abi.KNN_RADIUS, knn=6)
    qv = native.QueryBatchView.to_device(w.queries)
    op = bre.evalBre(qv).cpu().numpy()
    ob = bm.evalBeamBeamEstimate(qv).cpu().numpy()
    sp, sb = native.hit_signatures(ctx, bre, bm, qv)
    # oracle built with the oracle's own KNN radii
    rph = po.photon_knn_radius(w.photons.position, 0.3, 10)
    ph = scenes.Photons(w.photons.position, w.photons.inc_dir, w.photons.power, rph, 0.0)
    r0, r1 = po.beam_knn_radius(w.beams, 0.05, 6)
    oph = po.OraclePhotons(ph)
    obm = po.OracleBeams(w.beams, grid_res=0, start_radius=r0, end_radius=r1)
    rp, rsp, _ = po.bre_eval(oph, w.media, w.queries, signatures=True)
    rb, rsb, _ = po.beams_eval(obm, w.media, w.queries, signatures=True)
    assert (rp.max(1) > 0).sum() > 500 and (rb.max(1) > 0).sum() > 500
    assert np.array_equal(sp, rsp) and np.array_equal(sb, rsb)
    assert rel_err(op, rp).max() <= REL_TOL and rel_err(ob, rb).max() <= REL_TOL
