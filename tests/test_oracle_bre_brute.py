"""The oracle's EmbreeBre::evalBreBruteForce (SmallUPBP/src/Bre/Bre.cxx:980-1029), the reference's
debug estimate, pinned by closed forms:
- one photon on the ray: flux exp(-sigma_t t) rho / r^2 * sigma_s / pi, with the attenuation
  from the ray origin (not from min_t);
- the kernel is constant over the disc (an off-axis photon inside the radius counts the same),
  zero outside it and outside (min_t, max_t);
- the phase is PhaseFunction::Evaluate(direction, incDir): Henyey-Greenstein of
  cos = -dot(incDir, direction), checked against the textbook formula for incDir = +-direction;
- photons add linearly."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes


def _media(g=0.0, sa=(0.1, 0.2, 0.3), ss=(0.5, 0.25, 0.125)):
    return scenes.make_media([(sa, ss, g, 1.0)])


def _photons(pos, inc, power, r):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    return scenes.Photons(pos, np.asarray(inc, np.float32).reshape(-1, 3), np.asarray(power, np.float32).reshape(-1, 3),
                          None, float(r))


def _eval(ph, media, o, d, t0, t1):
    oph = po.OraclePhotons(ph, build_bvh=False)
    return po.bre_eval_brute_force(oph, media, [o], [d], [t0], [t1], [0])[0].astype(np.float64)


def _hg(g, cos_t):
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_t) ** 1.5)


def test_single_photon_closed_form():
    media = _media()
    st = np.array([0.6, 0.45, 0.425])
    ss = np.array([0.5, 0.25, 0.125])
    r, t = 0.05, 0.8
    flux = np.array([2.0, 1.0, 0.5])
    ph = _photons([0.0, 0.0, t], [1.0, 0.0, 0.0], flux, r)
    got = _eval(ph, media, [0, 0, 0], [0, 0, 1], 0.5, 2.0)
    want = flux * np.exp(-st * t) * (1 / (4 * math.pi)) / (r * r) * ss / math.pi
    assert got == pytest.approx(want, rel=2e-6)


def test_constant_disc_kernel_and_interval():
    media = _media()
    r = 0.05
    on = _eval(_photons([0, 0, 0.8], [1, 0, 0], [1, 1, 1], r), media, [0, 0, 0], [0, 0, 1], 0.0, 2.0)
    inside = _eval(_photons([0.049, 0, 0.8], [1, 0, 0], [1, 1, 1], r), media, [0, 0, 0], [0, 0, 1], 0.0, 2.0)
    outside = _eval(_photons([0.051, 0, 0.8], [1, 0, 0], [1, 1, 1], r), media, [0, 0, 0], [0, 0, 1], 0.0, 2.0)
    before = _eval(_photons([0, 0, 0.8], [1, 0, 0], [1, 1, 1], r), media, [0, 0, 0], [0, 0, 1], 0.9, 2.0)
    after = _eval(_photons([0, 0, 0.8], [1, 0, 0], [1, 1, 1], r), media, [0, 0, 0], [0, 0, 1], 0.0, 0.7)
    assert on.min() > 0
    assert inside == pytest.approx(on, rel=1e-6)
    assert not outside.any() and not before.any() and not after.any()


@pytest.mark.parametrize("g", [0.5, -0.3])
def test_phase_sign(g):
    media = _media(g=g)
    d = [0.0, 0.0, 1.0]
    ph_same = _photons([0, 0, 0.8], d, [1, 1, 1], 0.05)
    ph_opp = _photons([0, 0, 0.8], [0, 0, -1], [1, 1, 1], 0.05)
    a = _eval(ph_same, media, [0, 0, 0], d, 0.0, 2.0)
    b = _eval(ph_opp, media, [0, 0, 0], d, 0.0, 2.0)
    # Evaluate(direction, incDir): cos = -dot(incDir, direction)
    assert a[0] / b[0] == pytest.approx(_hg(g, -1.0) / _hg(g, 1.0), rel=1e-5)
    iso = _eval(ph_same, _media(g=0.0), [0, 0, 0], d, 0.0, 2.0)
    assert a[0] / iso[0] == pytest.approx(_hg(g, -1.0) * 4 * math.pi, rel=1e-5)


def test_photons_add_linearly():
    rng = np.random.default_rng(7)
    media = _media(g=0.3)
    n = 200
    pos = np.column_stack([rng.uniform(-0.02, 0.02, n), rng.uniform(-0.02, 0.02, n), rng.uniform(0.1, 1.9, n)])
    inc = rng.normal(size=(n, 3))
    inc /= np.linalg.norm(inc, axis=1, keepdims=True)
    pw = rng.uniform(0.1, 1.0, (n, 3))
    total = _eval(_photons(pos, inc, pw, 0.03), media, [0, 0, 0], [0, 0, 1], 0.0, 2.0)
    parts = sum(_eval(_photons(pos[i], inc[i], pw[i], 0.03), media, [0, 0, 0], [0, 0, 1], 0.0, 2.0)
                for i in range(n))
    assert total.min() > 0
    assert total == pytest.approx(parts, rel=1e-5)
