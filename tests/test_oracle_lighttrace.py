"""The oracle's light sub-path tracer (or_trace_light_paths) pinned from first principles.

Every produced quantity is recomputed in float64 from the vertex positions alone, with the
definitions of UPBP.hxx:501-674,1592-1662,1783-1870,1908-1969: the ray pieces between
boundary crossings and their media, the distance-sampling pdfs of HomogeneousMedium
(Media.hxx:240-323), the phase-function pdfs, the MisData bookkeeping (forward area pdf
inverses, reverse pdfs completed by the next vertex, the ray-sampling fields) and the
throughput. A wrong factor, index or order in the tracer fails these checks."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes

INV4PI = 1.0 / (4.0 * math.pi)


def _medium_params(media):
    st = (np.abs(media["absorption"]) + np.abs(media["scattering"])).astype(np.float64)
    ss = np.abs(media["scattering"]).astype(np.float64)
    mn = np.array([min([c for c in row if c > 0], default=0.0) for row in st])
    return st, ss, mn


def _medium_at(sc, p):
    m = sc.params["global_medium"]
    for g in sc.regions:
        if g["shape"] == scenes.REGION_SPHERE:
            inside = ((p - g["p0"]) ** 2).sum() <= g["p1"][0] ** 2
        else:
            inside = np.all(p >= g["p0"]) and np.all(p <= g["p1"])
        if inside:
            m = g["medium_id"]
    return m


def _crossings(sc, o, d):
    ts = []
    for g in sc.regions:
        if g["shape"] == scenes.REGION_SPHERE:
            oc = o - g["p0"]
            b = oc @ d
            c = oc @ oc - g["p1"][0] ** 2
            disc = b * b - c
            if disc >= 0:
                ts += [-b - math.sqrt(disc), -b + math.sqrt(disc)]
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                a0 = (g["p0"] - o) / d
                a1 = (g["p1"] - o) / d
            tn, tf = np.nanmax(np.minimum(a0, a1)), np.nanmin(np.maximum(a0, a1))
            if tn <= tf:
                ts += [tn, tf]
    return sorted(t for t in ts if t > 0)


def _pieces(sc, o, d, t_end):
    """(t0, t1, medium) of the ray up to t_end (the scatter distance)."""
    cut = [t for t in _crossings(sc, o, d) if t < t_end]
    b = [0.0] + cut + [t_end]
    return [(b[i], b[i + 1], _medium_at(sc, o + d * 0.5 * (b[i] + b[i + 1]))) for i in range(len(b) - 1)]


def _phase(g, cos_t):
    if abs(g) < 1e-3:
        return INV4PI
    dd = 1 + g * g - 2 * g * cos_t
    return INV4PI * (1 - g * g) / (dd * math.sqrt(dd))


def _check_paths(sc, lp, n_check=400, rtol=1e-4):
    """Positions are float32, so a quantity recomputed from them carries a relative error of
    about 1e-7 / (segment length) per length it depends on: each check's tolerance adds that
    term, and the direction-dependent ones (phase pdfs, sines) skip segments under 0.02."""
    st, ss, mn = _medium_params(sc.media)
    g_of = sc.media["mean_cosine"].astype(np.float64)
    cont_of = sc.media["continuation_prob"].astype(np.float64)
    lv, pos, ends = lp["vertices"], lp["positions"].astype(np.float64), lp["path_ends"]
    nl = len(sc.lights)
    starts = np.concatenate([[0], ends[:-1]])
    checked = 0
    for p in range(len(ends)):
        a, b = starts[p], ends[p]
        if b - a < 2:
            continue
        checked += 1
        if checked > n_check:
            break
        v = lv[a:b]
        assert np.array_equal(v["path_length"], np.arange(b - a)) and np.all(v["path_idx"] == p)
        # vertex 0: the point light (GenerateLightSample, delta and finite)
        assert v[0]["mis"]["pdf_a_inv"] == np.float32(nl) and v[0]["mis"]["rev_pdf_a"] == 0
        assert v[0]["mis"]["flags"] == scenes.MIS_DELTA | scenes.MIS_ON_LIGHT_SOURCE
        light = [i for i in range(nl) if np.allclose(sc.lights[i]["position"], pos[a])]
        assert light, "vertex 0 sits on a light"
        thr = sc.lights[light[0]]["intensity"].astype(np.float64) / (INV4PI / nl)
        n = b - a
        seg = {}
        for k in range(1, n):  # the ray from vertex k-1 to vertex k
            x0, x1 = pos[a + k - 1], pos[a + k]
            dist = np.linalg.norm(x1 - x0)
            d = (x1 - x0) / dist
            pdf, rev, att = 1.0, 1.0, np.ones(3)
            pcs = _pieces(sc, x0, d, dist)
            for j, (t0, t1, m) in enumerate(pcs):
                L = t1 - t0
                if ss[m].max() > 0 and mn[m] > 0:
                    e = math.exp(-mn[m] * L)
                    pdf *= mn[m] * e if j == len(pcs) - 1 else e
                    rev *= mn[m] * e if (k > 1 and j == 0) else e
                att *= np.exp(-st[m] * L)
            seg[k] = (dist, d, pdf, rev, att, pcs[-1][2])
        last_pdf_w_inv = (1.0 / nl) / (INV4PI / nl)
        acc_tol = 0.0  # the throughput accumulates every segment's error
        lpwi_tol = 0.0
        for k in range(1, n):
            dist, d, pdf, rev, att, m = seg[k]
            tol = rtol + 4e-7 * (1.0 + st[m].max()) / dist
            acc_tol += tol
            assert ss[m].max() > 0, "a vertex lies in a scattering medium"
            mis = v[k]["mis"]
            # forward area pdf inverse and distance-sampling fields (UPBP.hxx:589-616)
            if lpwi_tol >= 0:
                np.testing.assert_allclose(mis["pdf_a_inv"], last_pdf_w_inv * dist * dist / pdf, rtol=tol + lpwi_tol)
            np.testing.assert_allclose(mis["ray_sample_pdf_inv"], 1.0 / pdf, rtol=tol)
            np.testing.assert_allclose(mis["ray_sample_pdfs_ratio"], 1.0 / mn[m], rtol=1e-6)
            assert v[k]["flags"] == scenes.LV_CONNECTABLE | scenes.LV_IN_MEDIUM
            # vertex k-1's reverse data, completed by this ray (UPBP.hxx:620-622)
            pm = v[k - 1]["mis"]
            np.testing.assert_allclose(pm["ray_sample_rev_pdf_inv"], 1.0 / rev, rtol=tol)
            if k == 1:
                assert pm["rev_pdf_a"] == 0 and pm["rev_pdf_a_without_bsdf"] == 0  # delta light
            else:
                np.testing.assert_allclose(pm["rev_pdf_a_without_bsdf"], rev / (dist * dist), rtol=tol)
            # throughput: attenuation over the distance pdf (UPBP.hxx:546)
            thr = thr * att / pdf
            np.testing.assert_allclose(v[k]["throughput"], thr, rtol=acc_tol)
            lpwi_tol = -1.0  # unknown until the next direction is checked below
            if k + 1 < n:
                # SampleScattering at vertex k (UPBP.hxx:1908-1969): sigma_s * phase / (phase * cont)
                # on the throughput, 1 / (phase * cont) as the next last_pdf_w_inv, and the
                # reverse phase pdf * cont completing vertex k-1's reverse pdf
                dn = seg[k + 1][1]
                cos_t = float(d @ dn)
                ph = _phase(g_of[m], cos_t) * cont_of[m]
                thr = thr * ss[m] / cont_of[m]
                last_pdf_w_inv = 1.0 / ph
                assert mis["cos_theta_out"] == 1.0 and mis["flags"] & scenes.MIS_SPECULAR == 0
                if min(dist, seg[k + 1][0]) > 0.02:
                    lpwi_tol = 1e-3
                    np.testing.assert_allclose(mis["sin_theta"], math.sqrt(max(0.0, 1 - cos_t * cos_t)), atol=1e-3)
                    if k >= 2:
                        np.testing.assert_allclose(pm["rev_pdf_a"], pm["rev_pdf_a_without_bsdf"] * ph, rtol=1e-3)
    assert checked > 50


@pytest.mark.parametrize("g", [0.0, 0.6])
def test_light_paths_match_first_principles(g):
    sc = scenes.light_scene(mean_cosine=g, n_lights=2, max_path_length=8,
                            media_rows=[((0.05, 0.1, 0.02), (1.2, 1.0, 0.8), 0.0, 0.9),
                                        ((0, 0, 0), (0, 0, 0), 0.0, 1.0),
                                        ((0.1, 0.05, 0.2), (3.0, 2.5, 2.0), 0.0, 0.8)])
    lp = po.trace_light_paths(sc, 3000)
    _check_paths(sc, lp)


def test_light_source_and_lengths():
    sc = scenes.light_scene(max_path_length=4)
    lp = po.trace_light_paths(sc, 2000)
    lv, ends = lp["vertices"], lp["path_ends"]
    assert np.all(np.diff(np.concatenate([[0], ends])) >= 1)  # every path has its light vertex
    assert lv["path_length"].max() <= 3  # pathLength + 2 > mMaxPathLength stops the walk
    assert np.all(lv[ends[:-1]]["path_length"] == 0)
    # max_path_length 1: the reference traces no light paths at all
    sc1 = scenes.light_scene(max_path_length=1)
    lp1 = po.trace_light_paths(sc1, 100)
    assert len(lp1["vertices"]) == 0 and np.all(lp1["path_ends"] == 0)


def _beam_geometry(sc, lp):
    lv, pos = lp["vertices"], lp["positions"]
    st, ss, mn = _medium_params(sc.media)
    mfp = np.where(st.max(1) > 0, 1.0 / np.maximum(st.max(1), 1e-30), np.inf)
    for k in range(len(lp["beam_length"])):
        o, d, m = lp["beam_origin"][k], lp["beam_direction"][k], lp["beam_medium"][k]
        assert ss[m].max() > 0 and mfp[m] > sc.params["bb1d_min_mfp"]
        src = lp["beam_mis"][k]["light_vertex"]
        # the beam lies on the ray leaving its origin light vertex
        w = o.astype(np.float64) - pos[src]
        assert np.linalg.norm(np.cross(w, d)) < 1e-4
        mid = o + d * (0.5 * min(lp["beam_length"][k], 1.0))
        assert _medium_at(sc, mid) == m


@pytest.mark.parametrize("pbt", [scenes.SHORT_BEAM, scenes.LONG_BEAM])
def test_beams_follow_the_light_rays(pbt):
    sc = scenes.light_scene(photon_beam_type=pbt, max_path_length=6, bb1d_min_mfp=0.1)
    lp = po.trace_light_paths(sc, 2000)
    assert len(lp["beam_length"]) > 500
    _beam_geometry(sc, lp)
    assert np.all(lp["beam_flags"] == pbt)
    st = (np.abs(sc.media["absorption"]) + np.abs(sc.media["scattering"])).max(1)
    mbl = np.where(st > 0, -math.log(1e-9) / np.maximum(st, 1e-30), np.inf).astype(np.float32)
    assert np.all(lp["beam_length"] <= mbl[lp["beam_medium"]] * (1 + 1e-6))
    # every beam ends in a medium; rays leaving the light source are not "origin in medium"
    rsf = lp["beam_mis"]["ray_sampling_flags"]
    assert np.all(rsf & 2)
    from_light = lp["vertices"][lp["beam_mis"]["light_vertex"]]["path_length"] == 0
    assert from_light.sum() > 100 and np.all(rsf[from_light] & 1 == 0) and np.any(rsf[~from_light] & 1)


def test_beam_path_count_and_pb2d_filter():
    sc = scenes.light_scene(beam_path_count=500, bb1d_min_mfp=0.1)
    lp = po.trace_light_paths(sc, 2000)
    owner = lp["vertices"][lp["beam_mis"]["light_vertex"]]["path_idx"]
    assert owner.max() < 500
    # with PB2D merging, media with mfp <= bb1d_min_mfp store no beams and no IN_MEDIUM_WITH_BEAMS
    sc2 = scenes.light_scene(bb1d_min_mfp=0.5)
    lp2 = po.trace_light_paths(sc2, 2000)
    assert not np.any(lp2["beam_medium"] == 2)  # the sphere's mfp is 1/2.05
    in_sphere = (np.linalg.norm(lp2["positions"] - 0.5, axis=1) < 0.3) & (lp2["vertices"]["path_length"] > 0)
    assert in_sphere.sum() > 20
    assert np.all(lp2["vertices"][in_sphere]["mis"]["flags"] & scenes.MIS_IN_MEDIUM_WITH_BEAMS == 0)


def test_first_scatter_distance_is_exponential():
    """In one homogeneous medium filling the box, the first scatter distance from the light
    is Exp(sigma_t) truncated at the box boundary (SampleRay): check the mean of the
    untruncated part against the closed form."""
    rows = [((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), 0.0, 1.0), ((0, 0, 0), (0, 0, 0), 0.0, 1.0),
            ((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), 0.0, 1.0)]
    sc = scenes.light_scene(media_rows=rows, max_path_length=2)
    sc.lights[0]["position"] = (0.5, 0.5, 0.5)
    lp = po.trace_light_paths(sc, 40000)
    lv, pos = lp["vertices"], lp["positions"]
    first = lv["path_length"] == 1
    r = np.linalg.norm(pos[first] - 0.5, axis=1)
    # the distance to the box boundary is at least 0.5: Exp(4) truncated at 0.5
    sig, c = 4.0, 0.5
    trunc_mean = (1 / sig - (c + 1 / sig) * math.exp(-sig * c)) / (1 - math.exp(-sig * c))
    inside = r[r < c]
    assert abs(inside.mean() - trunc_mean) < 4 * inside.std() / math.sqrt(len(inside))
    assert abs(len(inside) / first.sum() - 1.0) < 0.2  # most scatter within 0.5
