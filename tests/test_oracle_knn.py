"""CPU tests that pin the oracle's k-nearest-neighbour radii.

The reference reads the k-th nearest squared distance from Jensen's kd-tree
(SmallUPBP/src/Misc/KdTmpl.hxx:600-792) and turns it into radii
(photons: SmallUPBP/src/Bre/Bre.cxx:586-596; beams: SmallUPBP/src/Beams/PhBeams.cxx:92-110).
The oracle's bounded max-heap is checked against numpy's selection routine and
against lattice closed forms; the radius formulas against exact powers of two.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from smallupbp_amd import scenes


def _d2_matrix(p, q):
    d = p[None, :, :] - q[:, None, :]  # item - query, float32
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


@pytest.mark.parametrize("k", [1, 2, 3, 8, 16, 17, 64, 200])
def test_knn_heap_equals_numpy_selection(k):
    rng = np.random.default_rng(7 + k)
    p = rng.random((1500, 3), dtype=np.float32)
    q = np.concatenate([p[:700], rng.random((300, 3), dtype=np.float32) * 1.2 - 0.1]).astype(np.float32)
    got = po.knn_d2(p, q, k)
    ref = np.partition(_d2_matrix(p, q), k - 1, axis=1)[:, k - 1]
    assert np.array_equal(got, ref)
    # the query's own item counts: k = 1 at an item is distance 0
    if k == 1:
        assert np.all(got[:700] == 0)


def test_knn_with_fewer_items_than_k_is_the_farthest():
    rng = np.random.default_rng(3)
    p = rng.random((9, 3), dtype=np.float32)
    got = po.knn_d2(p, p, 20)
    assert np.array_equal(got, _d2_matrix(p, p).max(axis=1))


def _lattice(n, h):
    g = np.arange(n, dtype=np.float32) * np.float32(h)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)


@pytest.mark.parametrize("k,shell", [(2, 1), (7, 1), (8, 2), (19, 2), (20, 3), (27, 3), (28, 4)])
def test_knn_lattice_closed_form(k, shell):
    # spacing 1/8: coordinates and differences are exact, so the interior point's
    # neighbour shells are at h^2 (6 faces), 2h^2 (12 edges), 3h^2 (8 corners), 4h^2 (6 axis-2)
    h = 0.125
    p = _lattice(7, h)
    interior = np.all((p > 1.5 * h) & (p < 4.5 * h), axis=1)
    got = po.knn_d2(p, p[interior], k)
    assert np.all(got == np.float32(shell * h * h))


def test_photon_knn_radius_formula_and_filter():
    h = 0.125
    p = _lattice(7, h)
    n = len(p)
    inm = np.ones(n, np.uint8)
    r = po.photon_knn_radius(p, 0.5, 4, in_medium=inm)
    interior = np.all((p > 0.5 * h) & (p < 5.5 * h), axis=1)
    # radius * (2 sqrt(h^2)) = 0.5 * 0.25
    assert np.all(r[interior] == np.float32(0.125))
    # dropping every other photon along x: the nearest kept neighbours are then at
    # h (y/z faces, 4 of them), sqrt(2) h (y/z diagonals, 4) and 2h (x axis, 2), so with
    # the photon itself k = 5 is at h^2 and k = 6 at 2h^2
    keep = (np.round(p[:, 0] / h).astype(int) % 2 == 0).astype(np.uint8)
    r5 = po.photon_knn_radius(p, 1.0, 5, in_medium=keep)
    assert np.all(r5[keep == 0] == 0)
    inner = interior & (keep == 1)
    assert np.all(r5[inner] == np.float32(2 * h))
    r6 = po.photon_knn_radius(p, 1.0, 6, in_medium=keep)
    assert np.all(r6[inner] == np.float32(2 * np.sqrt(np.float32(2 * h * h))))


def test_beam_knn_radius_ends_and_clamp():
    # four beams on a line of origins 1/4 apart, pointing +z with length 1/2
    o = np.array([[0, 0, 0], [0.25, 0, 0], [0.5, 0, 0], [0.75, 0, 0]], np.float32)
    d = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    ln = np.full(4, 0.5, np.float32)
    bm = scenes.Beams(o, d, ln, np.ones((4, 3), np.float32), np.full(4, scenes.SHORT_BEAM, np.uint32),
                      np.zeros(4, np.int32), 0.01)
    r0, r1 = po.beam_knn_radius(bm, 1.0, 2)
    # start: the 2nd nearest origin (itself first) is 1/4 away -> 2 * 0.25
    assert np.all(r0 == np.float32(0.5))
    # end (x, 0, 1/2): 2nd nearest origin at sqrt(1/16 + 1/4)
    e = np.float32(0.0625) + np.float32(0.25)
    assert np.allclose(r1, 2 * np.sqrt(e), rtol=0, atol=0)
    # a tiny multiplier is clamped to SMALLEST_RADIUS
    r0s, r1s = po.beam_knn_radius(bm, 1e-4, 2)
    assert np.all(r0s == np.float32(0.001)) and np.all(r1s == np.float32(0.001))
