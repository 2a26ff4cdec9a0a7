"""CPU tests that pin the oracle (no GPU).

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned by what the mathematics fixes: closed forms, kernel
normalisation integrals, invariants and brute-force equivalence. Each check is
chosen so that a plausible slip (dropped term, wrong sign, strict vs inclusive
bound, transposed operand) fails it.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes

F = np.float32


def c3(*v):
    return np.ascontiguousarray(v, np.float32)


def lib():
    return po.lib()


# ---------------------------------------------------------------- exp, hash
def test_cephes_exp_accuracy_and_clamp():
    xs = np.linspace(-87.0, 10.0, 20001).astype(np.float32)
    got = np.array([lib().or_exp_cephes(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(got - ref) / ref) < 4e-7
    # clamp at -88.3762626647949 (sse_mathfun.h exp_lo), not at float underflow
    assert lib().or_exp_cephes(-200.0) == pytest.approx(math.exp(-88.3762626647949), rel=1e-6)
    assert lib().or_exp_cephes(0.0) == 1.0


def test_hash_matches_vectorised_copy():
    idx = np.arange(0, 100000, 7, dtype=np.uint32)
    a = np.array([lib().or_hash32(int(i)) for i in idx[:500]], np.uint32)
    assert np.array_equal(a, po.hash32(idx[:500]))
    assert len(np.unique(po.hash32(idx))) == len(idx)


# ---------------------------------------------------------------- media
def test_medium_min_component_rules():
    import ctypes as C

    from smallupbp_amd import _abi

    class OrMedium(C.Structure):
        _fields_ = [("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("sigma_t", C.c_float * 3),
                    ("mean_cosine", C.c_float), ("continuation_prob", C.c_float), ("min_comp", C.c_float),
                    ("min_idx", C.c_int32), ("has_scattering", C.c_int32), ("has_attenuation", C.c_int32),
                    ("mfp", C.c_float), ("max_beam_length", C.c_float)]

    L = lib()
    L.or_prepare_media.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    rows = [((0.1, 0.1, 0.1), (0.2, 0.2, 0.2), 0, 1),  # tie -> lowest index 0
            ((0.0, 0.0, 0.0), (0.0, 0.3, 0.2), 0, 1),  # zero component skipped -> index 2
            ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0, 1),  # clear medium: index 0, no scattering
            ((-0.1, 0.0, 0.0), (0.0, -0.5, 0.2), 0, 1)]  # absolute values are taken
    m = scenes.make_media(rows)
    out = (OrMedium * len(m))()
    L.or_prepare_media(m.ctypes.data, len(m), C.addressof(out))
    assert (out[0].min_idx, out[0].min_comp) == (0, pytest.approx(0.3))
    assert (out[1].min_idx, out[1].min_comp) == (2, pytest.approx(0.2))
    assert out[2].min_idx == 0 and out[2].has_scattering == 0 and out[2].has_attenuation == 0
    assert list(out[3].sigma_t) == pytest.approx([0.1, 0.5, 0.2])
    assert out[3].min_idx == 0
    assert out[0].max_beam_length == pytest.approx(-math.log(1e-9) / 0.3, rel=1e-6)


# ---------------------------------------------------------------- phase
def test_phase_isotropic_and_hg_normalisation():
    d = c3(0, 0, 1)
    ex, ey = c3(1, 0, 0), c3(0, 1, 0)
    assert lib().or_phase_pdf(d.ctypes.data, ex.ctypes.data, 0.0, None) == pytest.approx(1 / (4 * math.pi))
    assert lib().or_phase_pdf(d.ctypes.data, ey.ctypes.data, 5e-4, None) == pytest.approx(1 / (4 * math.pi))
    for g in (-0.7, 0.3, 0.9):
        # integrate over the sphere of generated directions: must be 1
        n = 20000
        mu = (np.arange(n) + 0.5) / n * 2 - 1
        tot = 0.0
        for m in mu:
            gen = c3(0, math.sqrt(max(0.0, 1 - m * m)), m)
            tot += lib().or_phase_pdf(d.ctypes.data, gen.ctypes.data, g, None) * (2.0 / n) * 2 * math.pi
        assert tot == pytest.approx(1.0, rel=2e-3)
    # cos_theta = -dot(gen, fix): forward scattering g > 0 peaks when gen = -fix
    mz, pz = c3(0, 0, -1), c3(0, 0, 1)
    fwd = lib().or_phase_pdf(d.ctypes.data, mz.ctypes.data, 0.8, None)
    bwd = lib().or_phase_pdf(d.ctypes.data, pz.ctypes.data, 0.8, None)
    assert fwd > 10 * bwd


# ---------------------------------------------------------------- predicates
def _bre(o, d, tmin, tmax, p, r2):
    t = C_f()
    s = C_f()
    oa, da, pa = c3(*o), c3(*d), c3(*p)  # keep the arrays alive across the call
    hit = lib().or_test_bre(oa.ctypes.data, da.ctypes.data, tmin, tmax, pa.ctypes.data, r2, po.C.byref(t),
                            po.C.byref(s))
    return hit, t.value, s.value


def C_f():
    return po.C.c_float()


def test_bre_predicate_closed_forms():
    o, d = (0, 0, 0), (0, 0, 1)
    assert _bre(o, d, 1.0, 3.0, (0, 0, 2), 0.01) == (1, 2.0, 0.0)
    assert _bre(o, d, 2.0, 3.0, (0, 0, 2), 0.01)[0] == 0  # t == tmin: strict
    assert _bre(o, d, 1.0, 2.0, (0, 0, 2), 0.01)[0] == 0  # t == tmax: strict
    assert _bre(o, d, 1.0, 3.0, (0.5, 0, 2), 0.25) == (1, 2.0, 0.25)  # s == r^2: inclusive
    assert _bre(o, d, 1.0, 3.0, (0.5, 0, 2), 0.2499)[0] == 0
    # perpendicular offset: t is the foot of the perpendicular, s the squared offset
    hit, t, s = _bre((0, 0, 0), (0.6, 0.8, 0), 0.0, 10.0, (3.0, 4.0, 1.0), 4.0)
    assert hit and t == pytest.approx(5.0) and s == pytest.approx(1.0)


def _bb(o1, d1, a, b, o2, d2, mn, mx, r2):
    dist, st, t1, t2 = C_f(), C_f(), C_f(), C_f()
    arrs = [c3(*o1), c3(*d1), c3(*o2), c3(*d2)]
    hit = lib().or_test_beam_beam(arrs[0].ctypes.data, arrs[1].ctypes.data, a, b, arrs[2].ctypes.data,
                                  arrs[3].ctypes.data, mn, mx, r2, po.C.byref(dist), po.C.byref(st),
                                  po.C.byref(t1), po.C.byref(t2))
    return hit, dist.value, st.value, t1.value, t2.value


def test_beam_predicate_closed_forms():
    # exactly perpendicular lines: d1.d2 = 0 makes t2 = (t1 + d1.O1 - d1.O2) / (d1.d2) = 0/0,
    # which the reference rejects through its isnan test (PhotonBeam.hxx:284-286)
    h = 0.01
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, (-1, h, 2), (1, 0, 0), 0.0, 3.0, 0.0004)[0] == 0
    # oblique beam 30 degrees off the query axis, crossing at height h above it:
    # closest points at t2 = 1 on the beam, t1 = 1 + sin 60 on the query, dist = h, sin(theta) = 0.5
    d2 = (0.5, 0.0, math.sqrt(3) / 2)
    o2 = (-0.5, h, 1.0)
    hit, dist, st, t1, t2 = _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, o2, d2, 0.0, 3.0, 0.0004)
    assert hit and dist == pytest.approx(h, rel=1e-4) and st == pytest.approx(0.5, rel=1e-6)
    assert t1 == pytest.approx(1 + math.sqrt(3) / 2, rel=1e-6) and t2 == pytest.approx(1.0, rel=1e-5)
    # too far apart: ad^2 >= r2max * s2
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, (-0.5, 0.03, 1.0), d2, 0.0, 3.0, 0.0004)[0] == 0
    # parallel and near-parallel lines are rejected (|q^2 - 1| < 1e-5)
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, (0.001, 0, 0), (0, 0, 1), 0.0, 3.0, 0.01)[0] == 0
    e = 1e-3
    dn = np.array([e, 0, 1.0]) / math.sqrt(1 + e * e)
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, (0.0, 0.001, 0), tuple(dn), 0.0, 3.0, 0.01)[0] == 0
    # strict ranges on both lines
    assert _bb((0, 0, 0), (0, 0, 1), 1.9, 5.0, o2, d2, 0.0, 3.0, 0.0004)[0] == 0
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 1.8, o2, d2, 0.0, 3.0, 0.0004)[0] == 0
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, o2, d2, 0.0, 0.99, 0.0004)[0] == 0
    assert _bb((0, 0, 0), (0, 0, 1), 0.0, 5.0, o2, d2, 1.01, 3.0, 0.0004)[0] == 0


# ---------------------------------------------------------------- estimator closed forms
def _one_photon_workload(p, power, radius, media, segs, beam_type=scenes.LONG_BEAM, inc=(0, 0, 1)):
    ph = scenes.Photons(c3(*p).reshape(1, 3), c3(*inc).reshape(1, 3), c3(*power).reshape(1, 3), None, radius)
    o = c3(0, 0, 0).reshape(1, 3)
    d = c3(0, 0, 1).reshape(1, 3)
    off = np.array([0, len(segs)], np.int64)
    seg = scenes.segment_values(media, np.array([s[2] for s in segs], np.int32), np.array([s[0] for s in segs]),
                                np.array([s[1] for s in segs]))
    return ph, scenes.Queries(o, d, off, seg, beam_type)


def test_bre_single_photon_on_axis():
    media = scenes.make_media([((0.1, 0.2, 0.3), (0.5, 0.4, 0.3), 0.0, 1.0)])
    ph, q = _one_photon_workload((0, 0, 2.5), (1.0, 2.0, 3.0), 0.1, media, [(1.0, 4.0, 0)])
    for precision in (32, 64):
        out, _, cnt = po.bre_eval(po.OraclePhotons(ph), media, q, accel=0, precision=precision)
        st = np.array([0.6, 0.6, 0.6])
        expect = np.array([1, 2, 3]) * np.array([0.5, 0.4, 0.3]) * np.exp(-st * 1.5) / (4 * math.pi) * 2 / (
            math.pi * 0.01)
        assert cnt[0] == 1
        # inputs are float32 (sigma_t is formed in float as in Media.hxx:140), so FP64 differs by ~1e-7
        assert out[0] == pytest.approx(expect, rel=1e-6)


def test_bre_short_query_grey_medium_removes_transmittance():
    media = scenes.make_media([((0.1, 0.1, 0.1), (0.5, 0.5, 0.5), 0.0, 1.0)])
    outs = []
    for z in (1.5, 2.5, 3.5):
        ph, q = _one_photon_workload((0, 0, z), (1, 1, 1), 0.1, media, [(1.0, 4.0, 0)], beam_type=scenes.SHORT_BEAM)
        outs.append(po.bre_eval(po.OraclePhotons(ph), media, q, accel=0, precision=64)[0][0])
    assert np.allclose(outs, outs[0], rtol=1e-12)
    assert outs[0][0] == pytest.approx(0.5 / (4 * math.pi) * 2 / (math.pi * 0.01), rel=1e-6)


def test_epanechnikov_2d_kernel_integrates_to_one():
    # photons on a lattice in the plane z = 2 with spacing hh; sum of kernel * cell area -> 1
    hh = 0.002
    r = 0.05
    g = np.arange(-40, 41) * hh
    X, Y = np.meshgrid(g, g)
    n = X.size
    pos = np.stack([X.ravel() + 0.0003, Y.ravel() - 0.0007, np.full(n, 2.0)], -1).astype(np.float32)
    power = np.ones((n, 3), np.float32)
    ph = scenes.Photons(pos, np.tile(c3(0, 0, 1), (n, 1)), power, None, r)
    media = scenes.make_media([((0.0, 0.0, 0.0), (1e-6, 1e-6, 1e-6), 0.0, 1.0)])
    q = scenes.Queries(c3(0, 0, 0).reshape(1, 3), c3(0, 0, 1).reshape(1, 3), np.array([0, 1], np.int64),
                       scenes.segment_values(media, np.array([0], np.int32), np.array([1.0]), np.array([3.0])))
    out = po.bre_eval(po.OraclePhotons(ph), media, q, accel=0, precision=64)[0][0]
    integral = out[0] / (1e-6 / (4 * math.pi)) * hh * hh
    assert integral == pytest.approx(1.0, rel=2e-3)


def test_segment_loop_prefix_skip_and_early_out():
    # photon in the 3rd segment: result = T(seg1) * T(seg2: non-scattering) * single-photon value
    media = scenes.make_media([((0.1, 0.1, 0.1), (0.5, 0.5, 0.5), 0.0, 1.0),
                               ((0.3, 0.3, 0.3), (0.0, 0.0, 0.0), 0.0, 1.0)])
    ph, q = _one_photon_workload((0, 0, 3.0), (1, 1, 1), 0.1, media, [(0.0, 1.0, 0), (1.0, 2.0, 1), (2.0, 4.0, 0)])
    out = po.bre_eval(po.OraclePhotons(ph), media, q, accel=0, precision=64)[0][0]
    single = 0.5 * math.exp(-0.6 * 1.0) / (4 * math.pi) * 2 / (math.pi * 0.01)
    assert out[0] == pytest.approx(math.exp(-0.6) * math.exp(-0.3) * single, rel=1e-6)
    # a segment whose attenuation is below 1e-20 stops the loop before the photon's segment
    ph, q = _one_photon_workload((0, 0, 3.0), (1, 1, 1), 0.1, media, [(0.0, 1.0, 0), (1.0, 200.0, 1),
                                                                       (200.0, 400.0, 0)])
    ph.position[0, 2] = 300.0
    out = po.bre_eval(po.OraclePhotons(ph), media, q, accel=0, precision=64)[0][0]
    assert np.all(out == 0)


def test_beam_kernel_1d_integrates_to_one_and_on_axis_value():
    r = 0.02
    s3 = math.sqrt(3) / 2
    xs = (np.arange(-60, 61) * 0.0005 + 0.00013).astype(np.float32)
    n = len(xs)
    # beams 30 degrees off the query axis (sin theta = 0.5; an exactly perpendicular beam is
    # rejected by the reference's t2 formula), crossing the query at height y = x_i
    o = np.stack([np.full(n, -0.5), xs, np.full(n, 1.0)], -1).astype(np.float32)
    d = np.tile(c3(0.5, 0, s3), (n, 1))
    media = scenes.make_media([((0.0, 0.0, 0.0), (1e-6, 1e-6, 1e-6), 0.0, 1.0)])
    bm = scenes.Beams(o, d, np.full(n, 3.0, np.float32), np.ones((n, 3), np.float32),
                      np.full(n, scenes.LONG_BEAM, np.uint32), np.zeros(n, np.int32), r)
    q = scenes.Queries(c3(0, 0, 0).reshape(1, 3), c3(0, 0, 1).reshape(1, 3), np.array([0, 1], np.int64),
                       scenes.segment_values(media, np.array([0], np.int32), np.array([0.5]), np.array([3.0])))
    out = po.beams_eval(po.OracleBeams(bm, grid_res=0), media, q, accel=0, precision=64)[0][0]
    integral = out[0] / (1e-6 / (4 * math.pi)) * 0.0005 * 0.5
    assert integral == pytest.approx(1.0, rel=2e-3)
    # one beam through the query axis: Phi * sigma_s * T(t1 - tmin) * T(t2) * rho * 3 / (4 r sin theta)
    media = scenes.make_media([((0.1, 0.2, 0.05), (0.3, 0.3, 0.3), 0.0, 1.0)])
    bm1 = scenes.Beams(c3(-0.5, 0, 0.5).reshape(1, 3), c3(0.5, 0, s3).reshape(1, 3), np.array([3.0], np.float32),
                       c3(1, 2, 3).reshape(1, 3), np.array([scenes.LONG_BEAM], np.uint32), np.zeros(1, np.int32), r)
    q = scenes.Queries(c3(0, 0, 0).reshape(1, 3), c3(0, 0, 1).reshape(1, 3), np.array([0, 1], np.int64),
                       scenes.segment_values(media, np.array([0], np.int32), np.array([0.5]), np.array([3.0])))
    out = po.beams_eval(po.OracleBeams(bm1, grid_res=0), media, q, accel=0, precision=64)[0][0]
    st = np.array([0.4, 0.5, 0.35])
    expect = np.array([1, 2, 3]) * 0.3 * np.exp(-st * s3) * np.exp(-st * 1.0) / (4 * math.pi) * 3 / (4 * r * 0.5)
    assert out == pytest.approx(expect, rel=1e-5)
    # SHORT beam: beam transmittance divided by its minimum-sigma_t channel (index 2 here)
    bm1.flags[:] = scenes.SHORT_BEAM
    out = po.beams_eval(po.OracleBeams(bm1, grid_res=0), media, q, accel=0, precision=64)[0][0]
    assert out == pytest.approx(expect / math.exp(-0.35 * 1.0), rel=1e-5)
    # medium mismatch: a beam in another medium contributes nothing
    bm1.medium_id[:] = 1
    media2 = np.concatenate([media, media])
    out = po.beams_eval(po.OracleBeams(bm1, grid_res=0), media2, q, accel=0, precision=64)[0][0]
    assert np.all(out == 0)


# ---------------------------------------------------------------- accelerations vs brute force
@pytest.mark.parametrize("seed", [31, 32])
def test_bvh_hit_sets_equal_brute_force(seed):
    w = scenes.bre_beams(40, 24, 20_000, 10, seed=seed)
    oph = po.OraclePhotons(w.photons, leaf_size=3)
    a, sa, ca = po.bre_eval(oph, w.media, w.queries, accel=0, signatures=True)
    b, sb, cb = po.bre_eval(oph, w.media, w.queries, accel=1, signatures=True)
    assert np.array_equal(sa, sb)
    assert ca[0] == cb[0] and cb[1] < ca[1] / 10
    assert np.allclose(a, b, rtol=1e-5, atol=0)


def test_grid_matches_brute_force_except_boundary_ties():
    w = scenes.beams_stress(n_queries=1500, n_beams=6000, seed=33)
    obm = po.OracleBeams(w.beams, grid_res=64)
    a, sa, ca = po.beams_eval(obm, w.media, w.queries, accel=0, signatures=True)
    b, sb, cb = po.beams_eval(obm, w.media, w.queries, accel=1, signatures=True)
    mism = np.any(sa != sb, axis=1)
    # the grid drops hits whose closest point lies exactly on a cell boundary (strict cell ranges)
    assert mism.mean() < 0.005
    assert cb[0] <= ca[0]


def test_fp32_against_fp64_reevaluation():
    w = scenes.bre_beams(48, 27, 10_000, 3_000, seed=34)
    o32, s32, _ = po.bre_eval(po.OraclePhotons(w.photons), w.media, w.queries, accel=1, signatures=True)
    o64, s64, _ = po.bre_eval(po.OraclePhotons(w.photons), w.media, w.queries, accel=1, precision=64,
                              signatures=True)
    e = np.abs(o32 - o64).max(1) / np.maximum(np.abs(o64).max(1), 1e-30)
    assert np.median(e) < 1e-5
    assert np.mean(np.any(s32 != s64, axis=1)) < 0.01


# ---------------------------------------------------------------- adjoint (FP64) vs finite differences
def test_oracle_adjoint_matches_central_differences():
    w = scenes.adjoint_scene(12, 12, 3_000, 600, seed=35, chromatic=True)
    rng = np.random.default_rng(1)
    wp = rng.uniform(-1, 1, (w.queries.count, 3))
    wb = rng.uniform(-1, 1, (w.queries.count, 3))
    oph = po.OraclePhotons(w.photons, build_bvh=False)
    obm = po.OracleBeams(w.beams, grid_res=0)
    g = po.adjoint(oph, obm, w.media, w.queries, wp, wb)

    def loss(media, power=None):
        ww = scenes.with_media(w, media)
        ph = w.photons if power is None else scenes.Photons(w.photons.position, w.photons.inc_dir, power, None,
                                                            w.photons.const_radius)
        if w.photons.radius is not None and power is not None:
            ph.radius = w.photons.radius
        a = po.bre_eval(po.OraclePhotons(ph, build_bvh=False), media, ww.queries, accel=0, precision=64)[0]
        b = po.beams_eval(obm, media, ww.queries, accel=0, precision=64)[0]
        return float((wp * a).sum() + (wb * b).sum())

    h = 1e-3
    for m, c, field in ((0, 1, "scattering"), (2, 2, "absorption"), (2, 0, "scattering")):
        mp, mm = w.media.copy(), w.media.copy()
        mp[m][field][c] += h
        mm[m][field][c] -= h
        fd = (loss(mp) - loss(mm)) / (2 * h)
        an = g["grad_sigma_s" if field == "scattering" else "grad_sigma_a"][m, c]
        assert fd == pytest.approx(an, rel=1e-3, abs=1e-9 * max(1.0, abs(an)))
    # photon power is linear: perturb one photon with a hit
    k = int(np.argmax(np.abs(g["grad_photon_power"]).sum(1)))
    pp, pm = w.photons.power.copy(), w.photons.power.copy()
    pp[k, 1] += 1e-3
    pm[k, 1] -= 1e-3
    fd = (loss(w.media, pp) - loss(w.media, pm)) / 2e-3
    assert fd == pytest.approx(g["grad_photon_power"][k, 1], rel=1e-5)


@pytest.mark.parametrize("reduced", [False, True])
def test_oracle_adjoint_short_queries_match_central_differences(reduced):
    """SHORT query beams (Bre.cxx:257-263, PhotonBeam.hxx:336-345): in-segment transmittance
    T_c / T_k and, across segments, attenuation over the pass-through pdf exp(-sigma_t_k len), with
    k the minimum-sigma_t channel (unique in these chromatic media, so central differences stay on
    one side of its switch). reduced: the beam grid keeps 2 beams per crowded cell, each hit
    weighted by its count over the cell's pdf; the weights do not depend on sigma.
    The FP64 forward reads float32 segment records (attenuation, pdf), whose rounding puts noise of
    about 1e-7 L / h into central differences: at h = 1e-3 that exceeds 1e-3 of the smallest of the
    12 entries here, at h = 1e-2 it is below 1e-5 while the truncation error stays below 2e-4."""
    w = scenes.adjoint_scene(12, 12, 3_000, 600, seed=37, chromatic=True)
    q = w.queries
    w = scenes.with_media(dataclasses.replace(w, queries=scenes.Queries(q.origin, q.direction, q.seg_offsets,
                                                                      q.segments, scenes.SHORT_BEAM)), w.media)
    assert w.queries.beam_type == scenes.SHORT_BEAM
    rng = np.random.default_rng(4)
    wp = rng.uniform(-1, 1, (w.queries.count, 3))
    wb = rng.uniform(-1, 1, (w.queries.count, 3))
    oph = po.OraclePhotons(w.photons, build_bvh=False)
    if reduced:
        obm = po.OracleBeams(w.beams, grid_res=8).set_reduction(2, scenes.REDUCE_RESAMPLE_FIXED, 3)
    else:
        obm = po.OracleBeams(w.beams, grid_res=0)
    g = po.adjoint(oph, obm, w.media, w.queries, wp, wb)

    def loss(media):
        ww = scenes.with_media(w, media)
        a = po.bre_eval(oph, media, ww.queries, accel=0, precision=64)[0]
        b = po.beams_eval(obm, media, ww.queries, accel=0, precision=64)[0]
        return float((wp * a).sum() + (wb * b).sum())

    h = 1e-2
    for m in (0, 2):
        for field in ("scattering", "absorption"):
            for c in range(3):
                mp, mm = w.media.copy(), w.media.copy()
                mp[m][field][c] += h
                mm[m][field][c] -= h
                fd = (loss(mp) - loss(mm)) / (2 * h)
                an = g["grad_sigma_s" if field == "scattering" else "grad_sigma_a"][m, c]
                assert fd == pytest.approx(an, rel=1e-3, abs=1e-9 * max(1.0, abs(an))), (m, field, c)


def test_oracle_adjoint_carries_the_sign_of_negative_coefficients():
    """The media use |sigma| (Media.hxx:136-140), so d loss / d sigma for a negative input
    coefficient is sign(sigma) times the gradient with respect to |sigma|: central
    differences of the FP64 forward in the caller's (negative) value pin it."""
    w = scenes.adjoint_scene(12, 12, 3_000, 600, seed=36, chromatic=True)
    media = w.media.copy()
    media[0]["absorption"][1] = -media[0]["absorption"][1]
    media[2]["scattering"][0] = -media[2]["scattering"][0]
    w = scenes.with_media(w, media)
    rng = np.random.default_rng(2)
    wp = rng.uniform(-1, 1, (w.queries.count, 3))
    wb = rng.uniform(-1, 1, (w.queries.count, 3))
    oph = po.OraclePhotons(w.photons, build_bvh=False)
    obm = po.OracleBeams(w.beams, grid_res=0)
    g = po.adjoint(oph, obm, media, w.queries, wp, wb)

    def loss(m):
        ww = scenes.with_media(w, m)
        a = po.bre_eval(oph, m, ww.queries, accel=0, precision=64)[0]
        b = po.beams_eval(obm, m, ww.queries, accel=0, precision=64)[0]
        return float((wp * a).sum() + (wb * b).sum())

    h = 1e-3
    for m, c, field in ((0, 1, "absorption"), (2, 0, "scattering"), (0, 1, "scattering")):
        mp, mm = media.copy(), media.copy()
        mp[m][field][c] += h
        mm[m][field][c] -= h
        fd = (loss(mp) - loss(mm)) / (2 * h)
        an = g["grad_sigma_s" if field == "scattering" else "grad_sigma_a"][m, c]
        assert abs(an) > 0
        assert fd == pytest.approx(an, rel=1e-3, abs=1e-9 * max(1.0, abs(an)))
