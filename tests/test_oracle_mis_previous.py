"""Known answers for the PREVIOUS and BB1D_PREVIOUS approximations of the path-weight walks.

These modes drop terms on purpose, so their weights are not a balance heuristic and the
first-principles sums of test_oracle_mis_balance.py do not cover them. Here explicit short
paths carry MisData chosen so that every term of the walk is a distinct number; the expected
weights below are those numbers added up by hand from what the reference does
(SmallUPBP/src/Path/PathWeight.hxx):
  - camera walk, BB1D_PREVIOUS: the running BB1D sum is reset at every non-specular camera
    vertex (:61-62), so only the BB1D terms from the last non-specular vertex on (towards the
    camera) survive, while the other techniques' terms are kept;
  - light walk, BB1D_PREVIOUS: no BB1D term on the light side (:233, the walks here are never
    camera connections);
  - light walk, PREVIOUS: the walk stops after its first vertex (:266), i.e. after that
    vertex's BPT term.
Each case fails if the reset condition is inverted, the light-side skip dropped or the break
moved one step in either direction.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes

BPT, PB2D, BB1D = scenes.BPT, scenes.PB2D, scenes.BB1D
PREVIOUS, BB1D_PREVIOUS = scenes.PREVIOUS, scenes.BB1D_PREVIOUS
LONG = scenes.LONG_BEAM


def _mis(n):
    m = np.zeros(n, scenes.MIS_DTYPE)
    for f in ("pdf_a_inv", "rev_pdf_a", "rev_pdf_a_without_bsdf", "ray_sample_pdf_inv", "ray_sample_rev_pdf_inv",
              "ray_sample_pdfs_ratio", "ray_sample_rev_pdfs_ratio", "sin_theta", "cos_theta_out",
              "bb1d_beam_selection_pdf"):
        m[f] = 1.0
    m["flags"] = scenes.MIS_IN_MEDIUM_WITH_BEAMS
    return m


def _camera(spec1, spec2, spec3):
    """Camera path cam[0..3] (cam[3] is the merged vertex). The walk visits cam[3], cam[2],
    cam[1] with products 2, 6, 6 (last_rev 2; cam[3] forward 1/pdf 3; cam[2] reverse 0.5 times
    the next-to-last partial 2, forward 0.25; cam[1] reverse 4) and BB1D factors 1, 10, 100, so
    its BB1D terms are 2, 60 and 600; PB2D factors 1/8, 1/4, 1/2 give PB2D terms 0.25, 1.5, 3."""
    cam = _mis(4)
    cam["pdf_a_inv"][3], cam["rev_pdf_a"][2], cam["pdf_a_inv"][2], cam["rev_pdf_a"][1] = 3.0, 0.5, 0.25, 4.0
    cam["bb1d_factor"][1:4] = (100.0, 10.0, 1.0)
    cam["pb2d_factor"][1:4] = (0.5, 0.25, 0.125)
    for i, s in ((1, spec1), (2, spec2), (3, spec3)):
        if s:
            cam["flags"][i] |= scenes.MIS_SPECULAR
    return cam


def _cw(cam, tech):
    return po.camera_walk(cam, 3, 2.0, 2.0, tech, last_sin=1.0, last_ray_rev_inv=1.0, last_ray_rev_ratio=1.0,
                          qbt=LONG, pbt=LONG)


@pytest.mark.parametrize("spec,bb1d_kept", [
    ((True, True, False), 662.0),    # only cam[3] non-specular: reset before its term, all three kept
    ((True, True, True), 662.0),     # no non-specular vertex: nothing reset
    ((False, True, False), 600.0),   # cam[1] non-specular: reset there, only its term
    ((True, False, False), 660.0),   # cam[2] non-specular: reset there, cam[2] and cam[1] kept
    ((False, False, False), 600.0),  # every vertex non-specular: the last reset is at cam[1]
])
def test_camera_walk_bb1d_previous_resets_at_non_specular_vertices(spec, bb1d_kept):
    cam = _camera(*spec)
    assert _cw(cam, BB1D) == pytest.approx(662.0, rel=1e-12)
    assert _cw(cam, BB1D | BB1D_PREVIOUS) == pytest.approx(bb1d_kept, rel=1e-12)
    # other techniques' terms are not reset
    assert _cw(cam, BB1D | PB2D | BB1D_PREVIOUS) == pytest.approx(bb1d_kept + 4.75, rel=1e-12)


def _light_path():
    """One light path lv[0..3] (lv[0] on a point light, reverse pdf 0 ends the walk). The walk
    for a PB2D merge at lv[3] visits lv[3], lv[2], lv[1] with products 2, 6, 6 (last_rev 2;
    lv[3] forward 3; lv[2] reverse-without-BSDF 0.5 times the partial 2, forward 0.25; lv[1]
    reverse 4, forward 0.5) and stops at lv[0]. PB2D factors 1 and 100 at lv[2] and lv[1]
    give PB2D terms 6 and 600; BB1D factors 10 and 1000 give BB1D terms 60 and 6000. With BPT
    the terms after each forward step are 6, 1.5 and 3."""
    lv = np.zeros(4, scenes.LIGHT_VERTEX_DTYPE)
    lv["mis"] = _mis(4)
    lv["path_length"] = np.arange(4)
    lv["path_idx"] = 0
    lv["is_finite"] = 1
    m = lv["mis"]
    m["pdf_a_inv"][3], m["rev_pdf_a_without_bsdf"][2], m["pdf_a_inv"][2] = 3.0, 0.5, 0.25
    m["rev_pdf_a"][1], m["pdf_a_inv"][1], m["rev_pdf_a"][0] = 4.0, 0.5, 0.0
    m["pb2d_factor"][2], m["pb2d_factor"][1] = 1.0, 100.0
    m["bb1d_factor"][2], m["bb1d_factor"][1] = 10.0, 1000.0
    m["flags"][0] |= scenes.MIS_ON_LIGHT_SOURCE | scenes.MIS_DELTA
    return lv, np.array([4], np.int32)


def _lw(tech):
    lv, ends = _light_path()
    return po.light_walk(lv, ends, 0, 3, 2.0, 2.0, PB2D, tech, qbt=LONG, pbt=LONG)


def test_light_walk_bb1d_previous_drops_the_light_side_bb1d_terms():
    assert _lw(PB2D | BB1D) == pytest.approx(6.0 + 60.0 + 600.0 + 6000.0, rel=1e-12)
    assert _lw(PB2D | BB1D | BB1D_PREVIOUS) == pytest.approx(606.0, rel=1e-12)


def test_light_walk_previous_stops_after_the_first_vertex():
    assert _lw(PB2D | BPT) == pytest.approx(6.0 + 6.0 + 1.5 + 600.0 + 3.0, rel=1e-12)
    # the merged vertex's own merging terms are the camera side's; its BPT term is the only one kept
    assert _lw(PB2D | BPT | PREVIOUS) == pytest.approx(6.0, rel=1e-12)
    assert _lw(PB2D | BB1D | PREVIOUS) == 0.0
