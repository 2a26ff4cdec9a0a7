"""The oracle reproduces the committed golden fixtures bit for bit (CPU)."""
import numpy as np
import pytest

from oracle import pyoracle as po

from golden import fixtures

NAMES = fixtures.names()


def test_fixtures_present():
    assert len(NAMES) >= 4


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    w, z = fixtures.load(name)
    if w.photons is not None:
        o32, s32, _ = po.bre_eval(po.OraclePhotons(w.photons), w.media, w.queries, accel=1, signatures=True)
        assert np.array_equal(s32, z["sig_photons"])
        assert np.allclose(o32, z["out_photons_f32"], rtol=1e-6, atol=0)
        o32b, _, _ = po.bre_eval(po.OraclePhotons(w.photons, build_bvh=False), w.media, w.queries, accel=0)
        assert np.array_equal(o32b, z["out_photons_f32"])
        o64, _, _ = po.bre_eval(po.OraclePhotons(w.photons, build_bvh=False), w.media, w.queries, accel=0,
                                precision=64)
        assert np.array_equal(o64, z["out_photons_f64"])
    if w.beams is not None:
        obm = po.OracleBeams(w.beams, grid_res=0)
        o32, s32, _ = po.beams_eval(obm, w.media, w.queries, accel=0, signatures=True)
        assert np.array_equal(s32, z["sig_beams"])
        assert np.array_equal(o32, z["out_beams_f32"])
        o64, _, _ = po.beams_eval(obm, w.media, w.queries, accel=0, precision=64)
        assert np.array_equal(o64, z["out_beams_f64"])
