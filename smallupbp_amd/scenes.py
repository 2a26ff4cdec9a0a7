"""Seeded synthetic workloads for the volumetric merging estimators.

The reference produces photons, photon beams and query segments with its own
light and camera tracers (SmallUPBP/src/Renderers/VolLightTracer.hxx:168-271,586-640,
SmallUPBP/src/Renderers/UPBP.hxx:501-674,1783-1870, SmallUPBP/src/Scene/Scene.hxx:393-562).
Those producers are not part of the estimator path; here the same kinds of
inputs are generated analytically from a counter-based RNG (numpy Philox) so
the GPU library and the CPU oracle consume byte-identical arrays.

Geometry: an axis-aligned box medium, optionally with a sphere medium nested
inside it, surrounded by a clear (vacuum) medium. Query rays are LONG camera
beams; their volume segments carry what Scene::Intersect stores
(Scene.hxx:537-545): dist_min/dist_max, RaySamplePdf/RevPdf with no sampling
flags (Media.hxx:294-323) and attenuation exp(-sigma_t * length).

Workloads (sizes as named):
  plain_bre_256      unit cube, sigma_s=0.5, sigma_a=0.1, 100k photons r=0.01, 256x256 queries
  bre_beams_1080p    box + sphere, 1M photons (lognormal radii), 200k SHORT beams r=0.003, 1920x1080
  beams_stress       box sigma_s=1, sigma_a=0.1, 2M beams (50% SHORT / 50% LONG), 1M query beams
  batched_64         64 unit-box scenes of lognormal size (10k-2M photons, 16k-2M queries)
  adjoint_512        box + sphere, 500k photons + 100k beams, 512x512 queries
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

SHORT_BEAM = 1
LONG_BEAM = 2
# per-cell beam reduction types (UPBP_REDUCE_*)
REDUCE_PRESAMPLE, REDUCE_OFFSET, REDUCE_RESAMPLE_FIXED, REDUCE_RESAMPLE = 0, 1, 2, 3

MEDIUM_DTYPE = np.dtype(
    [
        ("absorption", "<f4", (3,)),
        ("scattering", "<f4", (3,)),
        ("mean_cosine", "<f4"),
        ("continuation_prob", "<f4"),
    ]
)
assert MEDIUM_DTYPE.itemsize == 32

SEGMENT_DTYPE = np.dtype(
    [
        ("dist_min", "<f4"),
        ("dist_max", "<f4"),
        ("ray_sample_pdf", "<f4"),
        ("ray_sample_rev_pdf", "<f4"),
        ("attenuation", "<f4", (3,)),
        ("medium_id", "<i4"),
    ]
)
assert SEGMENT_DTYPE.itemsize == 32


@dataclasses.dataclass
class Photons:
    position: np.ndarray  # [N,3] f32
    inc_dir: np.ndarray  # [N,3] f32
    power: np.ndarray  # [N,3] f32
    radius: Optional[np.ndarray]  # [N] f32 or None
    const_radius: float = 0.0

    @property
    def count(self) -> int:
        return int(self.position.shape[0])


@dataclasses.dataclass
class Beams:
    origin: np.ndarray  # [B,3] f32
    direction: np.ndarray  # [B,3] f32
    length: np.ndarray  # [B] f32
    throughput: np.ndarray  # [B,3] f32
    flags: np.ndarray  # [B] u32
    medium_id: np.ndarray  # [B] i32
    radius: float = 0.003

    @property
    def count(self) -> int:
        return int(self.origin.shape[0])


@dataclasses.dataclass
class Queries:
    origin: np.ndarray  # [Q,3] f32
    direction: np.ndarray  # [Q,3] f32
    seg_offsets: np.ndarray  # [Q+1] i64
    segments: np.ndarray  # [S] SEGMENT_DTYPE
    beam_type: int = LONG_BEAM
    lite: bool = False

    @property
    def count(self) -> int:
        return int(self.origin.shape[0])

    def subset(self, idx: np.ndarray) -> "Queries":
        """Queries `idx` (sorted or not) as a new batch with its own CSR."""
        idx = np.asarray(idx, dtype=np.int64)
        starts = self.seg_offsets[idx]
        counts = self.seg_offsets[idx + 1] - starts
        off = np.zeros(len(idx) + 1, np.int64)
        np.cumsum(counts, out=off[1:])
        gather = np.repeat(starts - off[:-1], counts) + np.arange(off[-1], dtype=np.int64)
        return Queries(
            np.ascontiguousarray(self.origin[idx]),
            np.ascontiguousarray(self.direction[idx]),
            off,
            np.ascontiguousarray(self.segments[gather]),
            self.beam_type,
            self.lite,
        )


@dataclasses.dataclass
class Workload:
    name: str
    media: np.ndarray  # MEDIUM_DTYPE
    photons: Optional[Photons]
    beams: Optional[Beams]
    queries: Queries
    sphere: Optional[tuple] = None  # (center[3], radius, medium_id)
    box_medium: int = 0
    vacuum_medium: int = 1
    seed: int = 0


def rng_for(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def make_media(rows) -> np.ndarray:
    """rows: (sigma_a[3], sigma_s[3], g, continuation_prob)."""
    m = np.zeros(len(rows), MEDIUM_DTYPE)
    for i, (sa, ss, g, cp) in enumerate(rows):
        m[i]["absorption"] = sa
        m[i]["scattering"] = ss
        m[i]["mean_cosine"] = g
        m[i]["continuation_prob"] = cp
    return m


def _sigma_t(media: np.ndarray) -> np.ndarray:
    return (np.abs(media["absorption"]) + np.abs(media["scattering"])).astype(np.float32)


def _min_positive(st: np.ndarray) -> np.ndarray:
    """Minimum positive component per medium, first index on ties (Media.hxx:142-145)."""
    out = np.zeros(st.shape[0], np.float32)
    for i in range(st.shape[0]):
        v, k = st[i, 0], 0
        for c in (1, 2):
            if st[i, c] > 0 and (v == 0 or st[i, c] < v):
                v, k = st[i, c], c
        out[i] = v
    return out


def _unit(v: np.ndarray) -> np.ndarray:
    n = np.sqrt((v.astype(np.float64) ** 2).sum(-1, keepdims=True))
    return (v / n).astype(np.float32)


def _isotropic_dirs(rng, n) -> np.ndarray:
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return _unit(np.stack([s * np.cos(phi), s * np.sin(phi), z], -1))


def _box_interval(o, d, lo, hi):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (lo - o) * inv
        t1 = (hi - o) * inv
    tmin = np.nanmax(np.minimum(t0, t1), axis=-1)
    tmax = np.nanmin(np.maximum(t0, t1), axis=-1)
    return tmin, tmax


def _sphere_interval(o, d, c, r):
    oc = o - c
    b = (d * oc).sum(-1)
    cc = (oc * oc).sum(-1) - r * r
    disc = b * b - cc
    hit = disc > 0
    sq = np.sqrt(np.where(hit, disc, 0.0))
    return hit, -b - sq, -b + sq


def segment_values(media: np.ndarray, medium_id: np.ndarray, tmin: np.ndarray, tmax: np.ndarray,
                   end_in_medium: Optional[np.ndarray] = None) -> np.ndarray:
    """Segment records as Scene::Intersect stores them for a ray that does not
    sample the medium (Scene.hxx:528-545): attenuation = exp(-sigma_t*(tmax-tmin)),
    pdf/revPdf = RaySamplePdf (Media.hxx:294-323)."""
    st = _sigma_t(media)
    mn = _min_positive(st)
    has_sc = (np.abs(media["scattering"]) > 0).any(-1)
    seg = np.zeros(len(tmin), SEGMENT_DTYPE)
    t0 = tmin.astype(np.float32)
    t1 = tmax.astype(np.float32)
    seg["dist_min"] = t0
    seg["dist_max"] = t1
    seg["medium_id"] = medium_id
    length = (t1 - t0).astype(np.float32)
    seg["attenuation"] = np.exp(-st[medium_id] * length[:, None]).astype(np.float32)
    att = np.maximum(np.exp(-mn[medium_id] * length), np.float32(1e-35)).astype(np.float32)
    samplable = (mn[medium_id] > 0) & has_sc[medium_id]
    pdf = np.where(samplable, att, np.float32(1.0)).astype(np.float32)
    if end_in_medium is not None:
        pdf = np.where(samplable & end_in_medium, (mn[medium_id] * att).astype(np.float32), pdf)
    seg["ray_sample_pdf"] = pdf
    seg["ray_sample_rev_pdf"] = np.where(samplable, att, np.float32(1.0)).astype(np.float32)
    return seg


def ray_segments(media, o, d, box_lo, box_hi, box_medium, vacuum_medium, sphere=None,
                 leading_vacuum=True, far=10.0):
    """Volume segments of rays through a box medium with an optional nested sphere."""
    o64 = o.astype(np.float64)
    d64 = d.astype(np.float64)
    n = o.shape[0]
    tin, tout = _box_interval(o64, d64, np.asarray(box_lo, np.float64), np.asarray(box_hi, np.float64))
    tin = np.maximum(tin, 0.0)
    hit_box = tout > tin
    if sphere is not None:
        hs, ts0, ts1 = _sphere_interval(o64, d64, np.asarray(sphere[0], np.float64), sphere[1])
        hs &= hit_box & (ts0 > tin) & (ts1 < tout)
    else:
        hs = np.zeros(n, bool)
    lead = leading_vacuum & hit_box & (tin > 0)
    counts = lead.astype(np.int64) + np.where(hit_box, np.where(hs, 3, 1), 0)
    counts = np.where(hit_box, counts, 1)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    total = int(off[-1])
    tmn = np.zeros(total)
    tmx = np.zeros(total)
    mid = np.zeros(total, np.int32)
    pos = off[:-1].copy()

    def put(mask, a, b, m):
        idx = pos[mask]
        tmn[idx] = a[mask]
        tmx[idx] = b[mask]
        mid[idx] = m
        pos[mask] += 1

    miss = ~hit_box
    put(miss, np.zeros(n), np.full(n, far), vacuum_medium)
    put(lead, np.zeros(n), tin, vacuum_medium)
    plain = hit_box & ~hs
    put(plain, tin, tout, box_medium)
    if sphere is not None:
        put(hs, tin, ts0, box_medium)
        put(hs, ts0, ts1, sphere[2])
        put(hs, ts1, tout, box_medium)
    assert np.array_equal(pos, off[1:])
    seg = segment_values(media, mid, tmn, tmx)
    return off, seg


def pinhole_rays(width, height, eye=(0.5, 0.5, -2.0)):
    """Pixel-centre rays from `eye` through the z=0 plane; the frame spans x in [0,1]
    and is centred at y = 0.5 with square pixels. Scanline order."""
    xs = (np.arange(width, dtype=np.float64) + 0.5) / width
    ys = 0.5 + ((np.arange(height, dtype=np.float64) + 0.5) - height * 0.5) / width
    X, Y = np.meshgrid(xs, ys)
    tgt = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], -1)
    e = np.asarray(eye, np.float64)
    d = tgt - e
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.broadcast_to(e, d.shape)
    return np.ascontiguousarray(o, dtype=np.float32), np.ascontiguousarray(d, dtype=np.float32)


def random_camera_rays(rng, n, eye=(0.5, 0.5, -2.0)):
    tgt = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.zeros(n)], -1)
    e = np.asarray(eye, np.float64)
    d = tgt - e
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.broadcast_to(e, d.shape)
    return np.ascontiguousarray(o, dtype=np.float32), np.ascontiguousarray(d, dtype=np.float32)


def _distance_to_boundary(o, d, sphere, box_lo=(0, 0, 0), box_hi=(1, 1, 1)):
    """Distance from points inside the box along d to the first medium boundary."""
    o64 = o.astype(np.float64)
    d64 = d.astype(np.float64)
    _, tout = _box_interval(o64, d64, np.asarray(box_lo, np.float64), np.asarray(box_hi, np.float64))
    dist = tout
    inside = np.zeros(o.shape[0], bool)
    if sphere is not None:
        c = np.asarray(sphere[0], np.float64)
        inside = ((o64 - c) ** 2).sum(-1) < sphere[1] ** 2
        hs, t0, t1 = _sphere_interval(o64, d64, c, sphere[1])
        # inside: exit at t1; outside: entry at t0 when ahead of the origin
        t_exit = np.where(inside, t1, np.where(hs & (t0 > 0), t0, np.inf))
        dist = np.minimum(dist, t_exit)
    return np.maximum(dist, 0.0), inside


def make_photons(rng, n, lo=(0, 0, 0), hi=(1, 1, 1), radius=None, radius_lognormal=None,
                 power_max=1e-4) -> Photons:
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    pos = (lo + rng.uniform(0, 1, (n, 3)) * (hi - lo)).astype(np.float32)
    inc = _isotropic_dirs(rng, n)
    power = rng.uniform(0, power_max, (n, 3)).astype(np.float32)
    if radius_lognormal is not None:
        mu, sigma, rmin, rmax = radius_lognormal
        r = np.clip(rng.lognormal(mu, sigma, n), rmin, rmax).astype(np.float32)
        return Photons(pos, inc, power, r, 0.0)
    return Photons(pos, inc, power, None, float(radius))


def make_beams(rng, n, media, sphere, box_medium, radius, short_fraction=1.0, short_mean=None,
               power_max=1e-4) -> Beams:
    """Beams with origin uniform in the unit box and isotropic direction.
    SHORT beams: length ~ Exp(mean) with mean = 1/min positive sigma_t of the origin medium
    (or `short_mean`), clipped at the first medium boundary. LONG beams: run to the boundary
    (UPBP.hxx:1798-1868)."""
    o = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    d = _isotropic_dirs(rng, n)
    bound, inside = _distance_to_boundary(o, d, sphere)
    med = np.full(n, box_medium, np.int32)
    if sphere is not None:
        med[inside] = sphere[2]
    st = _sigma_t(media)
    mn = _min_positive(st)
    is_short = rng.uniform(0, 1, n) < short_fraction
    mean = np.where(np.isfinite(1.0 / mn[med]), 1.0 / mn[med], 1.0) if short_mean is None else np.full(n, short_mean)
    exp_len = rng.exponential(1.0, n) * mean
    length = np.where(is_short, np.minimum(exp_len, bound), bound)
    length = np.maximum(length, 1e-4).astype(np.float32)
    flags = np.where(is_short, SHORT_BEAM, LONG_BEAM).astype(np.uint32)
    thr = rng.uniform(0, power_max, (n, 3)).astype(np.float32)
    return Beams(o, d, length, thr, flags, med, float(radius))


def _medium_rows_box_sphere():
    return [
        ((0.02, 0.02, 0.02), (0.2, 0.2, 0.2), 0.0, 0.8),  # 0: box
        ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 1.0),  # 1: vacuum (clear medium)
        ((0.05, 0.05, 0.05), (2.0, 2.0, 2.0), 0.0, 0.8),  # 2: sphere
    ]


SPHERE = ((0.5, 0.5, 0.5), 0.3, 2)


def plain_bre(width=256, height=256, n_photons=100_000, radius=0.01, seed=1, mean_cosine=0.0,
              beam_type=LONG_BEAM) -> Workload:
    """Plain BRE: unit cube, sigma_s=0.5, sigma_a=0.1, photons uniform, power U[0,1e-4], fixed radius."""
    rng = rng_for(seed)
    media = make_media([((0.1, 0.1, 0.1), (0.5, 0.5, 0.5), mean_cosine, 0.8),
                        ((0, 0, 0), (0, 0, 0), 0.0, 1.0)])
    ph = make_photons(rng, n_photons, radius=radius)
    o, d = pinhole_rays(width, height)
    off, seg = ray_segments(media, o, d, (0, 0, 0), (1, 1, 1), 0, 1)
    q = Queries(o, d, off, seg, beam_type)
    return Workload("plain_bre_%dx%d_%dph" % (width, height, n_photons), media, ph, None, q, None, 0, 1, seed)


def bre_beams(width=1920, height=1080, n_photons=1_000_000, n_beams=200_000, beam_radius=0.003,
              seed=2, mean_cosine=0.0) -> Workload:
    """Box + nested sphere, lognormal photon radii (mu=ln 0.005, sigma=0.4, clamped to
    [5e-4, 0.05]), SHORT beams with exponential lengths clipped at media boundaries."""
    rng = rng_for(seed)
    rows = _medium_rows_box_sphere()
    rows = [(a, s, mean_cosine if i != 1 else 0.0, c) for i, (a, s, g, c) in enumerate(rows)]
    media = make_media(rows)
    ph = make_photons(rng, n_photons, radius_lognormal=(math.log(0.005), 0.4, 5e-4, 0.05))
    bm = make_beams(rng, n_beams, media, SPHERE, 0, beam_radius, short_fraction=1.0)
    o, d = pinhole_rays(width, height)
    off, seg = ray_segments(media, o, d, (0, 0, 0), (1, 1, 1), 0, 1, SPHERE)
    q = Queries(o, d, off, seg, LONG_BEAM)
    return Workload("bre_beams_%dx%d_%dph_%dbm" % (width, height, n_photons, n_beams), media, ph, bm, q,
                    SPHERE, 0, 1, seed)


def beams_stress(n_queries=1_000_000, n_beams=2_000_000, beam_radius=0.003, seed=3) -> Workload:
    """One box medium sigma_s=1, sigma_a=0.1; 50% SHORT beams (Exp(mean 0.05) clipped to the box),
    50% LONG beams to the box boundary; query beams from random camera rays."""
    rng = rng_for(seed)
    media = make_media([((0.1, 0.1, 0.1), (1.0, 1.0, 1.0), 0.0, 0.8), ((0, 0, 0), (0, 0, 0), 0.0, 1.0)])
    bm = make_beams(rng, n_beams, media, None, 0, beam_radius, short_fraction=0.5, short_mean=0.05)
    o, d = random_camera_rays(rng, n_queries)
    off, seg = ray_segments(media, o, d, (0, 0, 0), (1, 1, 1), 0, 1)
    q = Queries(o, d, off, seg, LONG_BEAM)
    return Workload("beams_stress_%dq_%dbm" % (n_queries, n_beams), media, None, bm, q, None, 0, 1, seed)


CHROMATIC_ROWS = [
    ((0.02, 0.03, 0.01), (0.2, 0.25, 0.3), 0.0, 0.8),  # 0: box
    ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 1.0),  # 1: vacuum
    ((0.05, 0.04, 0.06), (2.0, 1.8, 2.2), 0.0, 0.8),  # 2: sphere
]


def adjoint_scene(width=512, height=512, n_photons=500_000, n_beams=100_000, beam_radius=0.003,
                  seed=5, scattering_scale=1.0, chromatic=False) -> Workload:
    """Box + sphere geometry with 500k photons + 100k beams at 512x512. `scattering_scale`
    multiplies sigma_s of both media (the target image uses 1.1). `chromatic` swaps in media
    whose minimum sigma_t channel is unique, so finite differences do not straddle the switch of
    the SHORT-beam normalisation channel (Media.hxx:142-145) that grey media sit on."""
    w = bre_beams(width, height, n_photons, n_beams, beam_radius, seed)
    if chromatic:
        w = with_media(w, make_media(CHROMATIC_ROWS))
    if scattering_scale != 1.0:
        w = with_media(w, scale_scattering(w.media, scattering_scale))
    return w


def scale_scattering(media: np.ndarray, s: float) -> np.ndarray:
    m = media.copy()
    m["scattering"] = (m["scattering"] * np.float32(s)).astype(np.float32)
    return m


def with_media(w: Workload, media: np.ndarray) -> Workload:
    """Same photons, beams and rays; segment attenuations and pdfs recomputed for `media`."""
    q = w.queries
    seg = segment_values(media, q.segments["medium_id"], q.segments["dist_min"], q.segments["dist_max"])
    q2 = Queries(q.origin, q.direction, q.seg_offsets, seg, q.beam_type, q.lite)
    return dataclasses.replace(w, media=media, queries=q2)


def batched_instances(n_instances=64, seed=4, scale=1.0, max_queries=2_000_000, workers=8) -> List[Workload]:
    """Independent unit-box scenes: sigma_s, sigma_a ~ U[0.05, 3] per channel; photons
    N ~ lognormal(mu=ln 1.4e5, sigma=1) clamped to [1e4, 2e6]; beams N/5; queries ~ U[16k, 2M]
    random camera rays. `scale` shrinks every count (tests use small scales). Instances are
    generated on `workers` threads; each has its own seeded generator, so the result does not
    depend on the thread count."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        rng = rng_for(seed * 1000 + i)
        sa = rng.uniform(0.05, 3.0, 3)
        ss = rng.uniform(0.05, 3.0, 3)
        media = make_media([(tuple(sa), tuple(ss), 0.0, 0.8), ((0, 0, 0), (0, 0, 0), 0.0, 1.0)])
        n_ph = int(np.clip(rng.lognormal(math.log(1.4e5), 1.0), 1e4, 2e6) * scale)
        n_ph = max(n_ph, 1)
        n_q = int(rng.uniform(16_000, max_queries) * scale)
        n_q = max(n_q, 1)
        ph = make_photons(rng, n_ph, radius_lognormal=(math.log(0.005), 0.4, 5e-4, 0.05))
        bm = make_beams(rng, max(n_ph // 5, 1), media, None, 0, 0.003, short_fraction=1.0)
        o, d = random_camera_rays(rng, n_q)
        off, seg = ray_segments(media, o, d, (0, 0, 0), (1, 1, 1), 0, 1)
        q = Queries(o, d, off, seg, LONG_BEAM)
        return Workload("instance_%02d" % i, media, ph, bm, q, None, 0, 1, seed * 1000 + i)

    if workers <= 1 or n_instances <= 1:
        return [one(i) for i in range(n_instances)]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(n_instances)))


# ---------------------------------------------------------------------------
# UPBP mode inputs (MIS data). Layouts match upbp_mis_data / upbp_light_vertex /
# upbp_beam_mis / upbp_query_mis of include/upbp.h.
MIS_DTYPE = np.dtype(
    [("pdf_a_inv", "<f4"), ("rev_pdf_a", "<f4"), ("rev_pdf_a_without_bsdf", "<f4"), ("ray_sample_pdf_inv", "<f4"),
     ("ray_sample_rev_pdf_inv", "<f4"), ("ray_sample_pdfs_ratio", "<f4"), ("ray_sample_rev_pdfs_ratio", "<f4"),
     ("sin_theta", "<f4"), ("cos_theta_out", "<f4"), ("surf_factor", "<f4"), ("pp3d_factor", "<f4"),
     ("pb2d_factor", "<f4"), ("bb1d_factor", "<f4"), ("bb1d_beam_selection_pdf", "<f4"), ("flags", "<u4"),
     ("pad", "<u4")])
assert MIS_DTYPE.itemsize == 64
LIGHT_VERTEX_DTYPE = np.dtype(
    [("throughput", "<f4", (3,)), ("dir_fix", "<f4", (3,)), ("continuation_prob", "<f4"), ("path_idx", "<i4"),
     ("path_length", "<u4"), ("is_finite", "<u4"), ("flags", "<u4"), ("pad", "<u4"), ("mis", MIS_DTYPE)])
assert LIGHT_VERTEX_DTYPE.itemsize == 112
BEAM_MIS_DTYPE = np.dtype([("light_vertex", "<i4"), ("ray_sampling_flags", "<u4"), ("ray_sample_pdf", "<f4"),
                           ("ray_sample_rev_pdf", "<f4"), ("last_pdf_w_inv", "<f4"), ("pad", "<u4", (3,))])
assert BEAM_MIS_DTYPE.itemsize == 32
QUERY_MIS_DTYPE = np.dtype([("camera_offset", "<i8"), ("camera_path_length", "<u4"), ("ray_sampling_flags", "<u4"),
                            ("last_pdf_w_inv", "<f4"), ("pad", "<u4")])
assert QUERY_MIS_DTYPE.itemsize == 24

PP3D_QUERY_DTYPE = np.dtype([("position", "<f4", (3,)), ("medium", "<i4"), ("dir_fix", "<f4", (3,)),
                             ("camera_path_length", "<u4"), ("camera_offset", "<i8"), ("pad", "<i8")])
assert PP3D_QUERY_DTYPE.itemsize == 48
SURF_QUERY_DTYPE = np.dtype([("position", "<f4", (3,)), ("camera_path_length", "<u4"), ("frame_x", "<f4", (3,)),
                             ("continuation_prob", "<f4"), ("frame_y", "<f4", (3,)), ("diff_prob", "<f4"),
                             ("frame_z", "<f4", (3,)), ("phong_prob", "<f4"), ("local_dir_fix", "<f4", (3,)),
                             ("phong_exponent", "<f4"), ("diffuse", "<f4", (3,)), ("pad0", "<f4"),
                             ("phong", "<f4", (3,)), ("pad1", "<f4"), ("camera_offset", "<i8"), ("pad2", "<i8")])
assert SURF_QUERY_DTYPE.itemsize == 128

MIS_DELTA, MIS_ON_LIGHT_SOURCE, MIS_SPECULAR, MIS_IN_MEDIUM_WITH_BEAMS = 1, 2, 4, 8
LV_CONNECTABLE, LV_IN_MEDIUM = 1, 2
VM_NO_MIS = 1
BPT, SURF, PP3D, PB2D, BB1D = 4, 8, 16, 32, 64
NO_SINE_IN_WEIGHTS, PREVIOUS, COMPATIBLE, BB1D_PREVIOUS = 128, 256, 512, 1024


@dataclasses.dataclass
class MisInputs:
    queries: np.ndarray  # QUERY_MIS_DTYPE [Q]
    camera: np.ndarray  # MIS_DTYPE
    light_vertices: np.ndarray  # LIGHT_VERTEX_DTYPE
    path_ends: np.ndarray  # int32
    beams: Optional[np.ndarray]  # BEAM_MIS_DTYPE [B]
    params: dict


def random_mis(rng, n, specular_p=0.1, beams_p=0.7) -> np.ndarray:
    m = np.zeros(n, MIS_DTYPE)
    for f in ("pdf_a_inv", "rev_pdf_a", "rev_pdf_a_without_bsdf", "ray_sample_pdfs_ratio", "ray_sample_rev_pdfs_ratio"):
        m[f] = rng.uniform(0.5, 2.0, n)
    m["ray_sample_pdf_inv"] = rng.uniform(1.0, 3.0, n)
    m["ray_sample_rev_pdf_inv"] = rng.uniform(1.0, 3.0, n)
    m["sin_theta"] = rng.uniform(0.2, 1.0, n)
    m["cos_theta_out"] = rng.uniform(-1.0, 1.0, n)
    for f in ("surf_factor", "pp3d_factor", "pb2d_factor", "bb1d_factor"):
        m[f] = rng.uniform(0.1, 2.0, n)
    m["bb1d_beam_selection_pdf"] = 1.0
    fl = np.where(rng.uniform(size=n) < specular_p, MIS_SPECULAR, 0)
    fl |= np.where(rng.uniform(size=n) < beams_p, MIS_IN_MEDIUM_WITH_BEAMS, 0)
    m["flags"] = fl.astype(np.uint32)
    return m


def upbp_scene(width=32, height=18, n_paths=2_000, n_beams=500, photon_radius=0.03, beam_radius=0.01, seed=41,
               techniques=PB2D | BB1D | BPT | PP3D | SURF, query_beam_type=LONG_BEAM, photon_beam_type=SHORT_BEAM,
               max_light_length=4, max_camera_length=3, media_rows=None):
    """Box + sphere media; light sub-paths of 1..max_light_length segments whose non-source
    vertices are photons in the media; beams leave random light vertices; each query has a
    camera sub-path of 1..max_camera_length vertices with random MIS data."""
    rng = rng_for(seed)
    media = make_media(media_rows or _medium_rows_box_sphere())
    lens = rng.integers(1, max_light_length + 1, n_paths)
    counts = lens + 1
    ends = np.cumsum(counts).astype(np.int32)
    nv = int(ends[-1])
    starts = np.concatenate([[0], ends[:-1]])
    lv = np.zeros(nv, LIGHT_VERTEX_DTYPE)
    path_idx = np.repeat(np.arange(n_paths), counts)
    path_len = np.arange(nv) - np.repeat(starts, counts)
    lv["path_idx"] = path_idx
    lv["path_length"] = path_len
    lv["throughput"] = rng.uniform(0, 1e-4, (nv, 3))
    lv["dir_fix"] = _isotropic_dirs(rng, nv)
    lv["continuation_prob"] = 0.8
    lv["is_finite"] = 1
    lv["mis"] = random_mis(rng, nv)
    src = path_len == 0
    # light-source vertices: on the light, no merging factors (they are not in a medium)
    lv["mis"]["flags"][src] = MIS_ON_LIGHT_SOURCE
    for f in ("surf_factor", "pp3d_factor", "pb2d_factor", "bb1d_factor"):
        lv["mis"][f][src] = 0.0
    pos = rng.uniform(0, 1, (nv, 3)).astype(np.float32)
    in_medium = (~src).astype(np.uint8)
    photons = Photons(pos, np.ascontiguousarray(lv["dir_fix"]), np.ascontiguousarray(lv["throughput"]), None,
                      photon_radius)
    bm = make_beams(rng, n_beams, media, SPHERE, 0, beam_radius, short_fraction=1.0 if photon_beam_type == SHORT_BEAM else 0.0)
    bmis = np.zeros(n_beams, BEAM_MIS_DTYPE)
    bmis["light_vertex"] = rng.integers(0, nv, n_beams)
    bmis["ray_sampling_flags"] = 2 | np.where(rng.uniform(size=n_beams) < 0.3, 1, 0)
    bmis["ray_sample_pdf"] = rng.uniform(0.5, 1.0, n_beams)
    bmis["ray_sample_rev_pdf"] = rng.uniform(0.5, 1.0, n_beams)
    bmis["last_pdf_w_inv"] = rng.uniform(0.5, 2.0, n_beams)
    o, d = pinhole_rays(width, height)
    off, seg = ray_segments(media, o, d, (0, 0, 0), (1, 1, 1), 0, 1, SPHERE)
    q = Queries(o, d, off, seg, query_beam_type)
    nq = q.count
    clen = rng.integers(1, max_camera_length + 1, nq)
    qm = np.zeros(nq, QUERY_MIS_DTYPE)
    coff = np.concatenate([[0], np.cumsum(clen + 1)[:-1]])
    qm["camera_offset"] = coff
    qm["camera_path_length"] = clen
    qm["ray_sampling_flags"] = np.where((clen > 1) & (rng.uniform(size=nq) < 0.5), 1, 0)
    qm["last_pdf_w_inv"] = rng.uniform(0.5, 2.0, nq)
    cam = random_mis(rng, int((clen + 1).sum()))
    params = dict(techniques=techniques, query_beam_type=query_beam_type, photon_beam_type=photon_beam_type,
                  min_path_length=0, max_path_length=20, surf_factor=0.7, pp3d_factor=0.9, pb2d_factor=0.6,
                  bb1d_factor=1.3, bb1d_min_mfp=0.5, min_dist_to_med=0.0, has_bb1d_beams=1)
    w = Workload("upbp_%dx%d" % (width, height), media, photons, bm, q, SPHERE, 0, 1, seed)
    return w, in_medium, MisInputs(qm, cam, lv, ends, bmis, params)


@dataclasses.dataclass
class MergeScene:
    """Inputs of PP3D (camera vertices in media) or SURF (camera vertices on surfaces) merging."""
    technique: int  # PP3D or SURF
    media: np.ndarray
    positions: np.ndarray  # [n][3] float32, light-vertex positions (mHitpoint)
    radius: float
    cell_count: int  # hash buckets (the reference reserves one per light path)
    queries: np.ndarray  # PP3D_QUERY_DTYPE or SURF_QUERY_DTYPE
    mis: MisInputs  # light vertices (with flags), camera MisData, path ends, params; queries unused


def medium_at(p: np.ndarray, sphere=SPHERE, box_medium=0, outside_medium=1) -> np.ndarray:
    """Medium id by position: the sphere inside the unit box, the box medium elsewhere in it."""
    c, r, sid = sphere
    inside_box = np.all((p >= 0) & (p <= 1), axis=-1)
    d2 = ((p.astype(np.float64) - np.asarray(c)) ** 2).sum(-1)
    return np.where(d2 <= r * r, sid, np.where(inside_box, box_medium, outside_medium)).astype(np.int32)


def _light_paths(rng, n_paths, max_light_length):
    """Light sub-paths of 1..max_light_length vertices after the source, with random MIS data;
    light-source vertices carry no merging factors."""
    lens = rng.integers(1, max_light_length + 1, n_paths)
    counts = lens + 1
    ends = np.cumsum(counts).astype(np.int32)
    nv = int(ends[-1])
    starts = np.concatenate([[0], ends[:-1]])
    lv = np.zeros(nv, LIGHT_VERTEX_DTYPE)
    path_len = np.arange(nv) - np.repeat(starts, counts)
    lv["path_idx"] = np.repeat(np.arange(n_paths), counts)
    lv["path_length"] = path_len
    lv["throughput"] = rng.uniform(0, 1e-3, (nv, 3))
    lv["dir_fix"] = _isotropic_dirs(rng, nv)
    lv["continuation_prob"] = rng.uniform(0.5, 1.0, nv)
    lv["is_finite"] = 1
    lv["mis"] = random_mis(rng, nv)
    src = path_len == 0
    lv["mis"]["flags"][src] = MIS_ON_LIGHT_SOURCE
    for f in ("surf_factor", "pp3d_factor", "pb2d_factor", "bb1d_factor"):
        lv["mis"][f][src] = 0.0
    return lv, ends, src


def _camera_paths(rng, n_queries, max_camera_length):
    clen = rng.integers(1, max_camera_length + 1, n_queries)
    offsets = np.concatenate([[0], np.cumsum(clen + 1)[:-1]])
    return clen, offsets, random_mis(rng, int((clen + 1).sum()))


def _params(techniques, query_beam_type, photon_beam_type):
    return dict(techniques=techniques, query_beam_type=query_beam_type, photon_beam_type=photon_beam_type,
                min_path_length=0, max_path_length=20, surf_factor=0.7, pp3d_factor=0.9, pb2d_factor=0.6,
                bb1d_factor=1.3, bb1d_min_mfp=0.5, min_dist_to_med=0.0, has_bb1d_beams=1)


def pp3d_scene(n_paths=2_000, n_queries=4_000, radius=0.05, cell_count=None, seed=61, max_light_length=4,
               max_camera_length=3, techniques=PB2D | BB1D | BPT | PP3D | SURF, query_beam_type=LONG_BEAM,
               photon_beam_type=SHORT_BEAM, mean_cosine=0.0, connectable_p=0.9, query_margin=0.05) -> MergeScene:
    """Light sub-paths whose non-source vertices lie in the box + sphere media (UPBP_LV_IN_MEDIUM), most of
    them connectable; camera vertices uniform in a box slightly larger than the unit box (so some lie
    outside the grid's bounding box), each with a camera sub-path of random MIS data."""
    rng = rng_for(seed)
    rows = _medium_rows_box_sphere()
    if mean_cosine:
        rows = [(sa, ss, mean_cosine if i != 1 else 0.0, cp) for i, (sa, ss, _, cp) in enumerate(rows)]
    media = make_media(rows)
    lv, ends, src = _light_paths(rng, n_paths, max_light_length)
    nv = len(lv)
    conn = rng.uniform(size=nv) < connectable_p
    lv["flags"] = np.where(~src, LV_IN_MEDIUM, 0) | np.where(conn & ~src, LV_CONNECTABLE, 0)
    pos = rng.uniform(0, 1, (nv, 3)).astype(np.float32)
    qp = rng.uniform(-query_margin, 1 + query_margin, (n_queries, 3)).astype(np.float32)
    clen, coff, cam = _camera_paths(rng, n_queries, max_camera_length)
    q = np.zeros(n_queries, PP3D_QUERY_DTYPE)
    q["position"] = qp
    q["medium"] = medium_at(qp)
    q["dir_fix"] = _isotropic_dirs(rng, n_queries)
    q["camera_path_length"] = clen
    q["camera_offset"] = coff
    mi = MisInputs(None, cam, lv, ends, None, _params(techniques, query_beam_type, photon_beam_type))
    return MergeScene(PP3D, media, pos, float(radius), int(cell_count or n_paths), q, mi)


# the three planes of the SURF scene: (axis of the normal, its sign, the plane's coordinate)
_SURF_PLANES = ((1, 1.0, 0.0), (2, -1.0, 1.0), (0, 1.0, 0.0))


def _on_planes(rng, n):
    """Points uniform on three faces of the unit box (floor y = 0, back wall z = 1, left wall x = 0),
    and the face normals."""
    k = rng.integers(0, len(_SURF_PLANES), n)
    p = rng.uniform(0, 1, (n, 3))
    nrm = np.zeros((n, 3))
    for i, (ax, sgn, c) in enumerate(_SURF_PLANES):
        m = k == i
        p[m, ax] = c
        nrm[m, ax] = sgn
    return p.astype(np.float32), nrm


def _hemisphere(rng, nrm, flip_p):
    d = _isotropic_dirs(rng, len(nrm)).astype(np.float64)
    side = np.sign((d * nrm).sum(-1))
    side[side == 0] = 1
    flip = rng.uniform(size=len(nrm)) < flip_p
    d *= np.where(flip, -side, side)[:, None]
    return d


def _frames(rng, nrm, tilt):
    """Orthonormal shading frames around the face normals tilted by up to `tilt` radians."""
    z = nrm + rng.normal(0, tilt, nrm.shape)
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    a = np.where(np.abs(z[:, :1]) > 0.99, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    y = np.cross(z, a)
    y /= np.linalg.norm(y, axis=1, keepdims=True)
    x = np.cross(y, z)
    return x, y, z


def surf_scene(n_paths=2_000, n_queries=4_000, radius=0.05, cell_count=None, seed=63, max_light_length=4,
               max_camera_length=3, techniques=PB2D | BB1D | BPT | PP3D | SURF, query_beam_type=LONG_BEAM,
               photon_beam_type=SHORT_BEAM, connectable_p=0.9, lobes="both", flip_p=0.08, tilt=0.05) -> MergeScene:
    """Light sub-paths whose non-source vertices lie on three faces of the unit box (no
    UPBP_LV_IN_MEDIUM), with incoming directions mostly on the face's side; camera vertices on the same
    faces with slightly tilted shading frames and diffuse + Phong materials whose lobe and continuation
    probabilities follow Bsdf::GetComponentProbabilities for opaque materials. lobes: "both",
    "diffuse" or "phong"."""
    rng = rng_for(seed)
    media = make_media(_medium_rows_box_sphere())
    lv, ends, src = _light_paths(rng, n_paths, max_light_length)
    nv = len(lv)
    conn = rng.uniform(size=nv) < connectable_p
    lv["flags"] = np.where(conn & ~src, LV_CONNECTABLE, 0)
    pos, lnrm = _on_planes(rng, nv)
    lv["dir_fix"] = _hemisphere(rng, lnrm, flip_p).astype(np.float32)
    qp, qn = _on_planes(rng, n_queries)
    x, y, z = _frames(rng, qn, tilt)
    fix_world = _hemisphere(rng, qn, flip_p)
    clen, coff, cam = _camera_paths(rng, n_queries, max_camera_length)
    q = np.zeros(n_queries, SURF_QUERY_DTYPE)
    q["position"] = qp
    q["frame_x"], q["frame_y"], q["frame_z"] = x, y, z
    q["local_dir_fix"] = np.stack([(fix_world * x).sum(-1), (fix_world * y).sum(-1), (fix_world * z).sum(-1)], -1)
    diff = rng.uniform(0.05, 0.8, (n_queries, 3)) * (lobes != "phong")
    phong = rng.uniform(0.05, 0.5, (n_queries, 3)) * (lobes != "diffuse")
    lum = np.array([0.212671, 0.715160, 0.072169])
    ad, ap = diff @ lum, phong @ lum
    q["diffuse"], q["phong"] = diff, phong
    q["phong_exponent"] = rng.uniform(1.0, 80.0, n_queries)
    q["diff_prob"] = ad / (ad + ap)
    q["phong_prob"] = ap / (ad + ap)
    q["continuation_prob"] = np.minimum(1.0, (diff + phong).max(axis=1))
    q["camera_path_length"] = clen
    q["camera_offset"] = coff
    mi = MisInputs(None, cam, lv, ends, None, _params(techniques, query_beam_type, photon_beam_type))
    return MergeScene(SURF, media, pos, float(radius), int(cell_count or n_paths), q, mi)


# ---- light sub-path tracing (upbp_trace_light_paths) ------------------------------------
REGION_SPHERE = 0
REGION_BOX = 1
REGION_DTYPE = np.dtype([("shape", "<u4"), ("medium_id", "<i4"), ("p0", "<f4", (3,)), ("p1", "<f4", (3,))])
POINT_LIGHT_DTYPE = np.dtype([("position", "<f4", (3,)), ("intensity", "<f4", (3,))])
assert REGION_DTYPE.itemsize == 32 and POINT_LIGHT_DTYPE.itemsize == 24


@dataclasses.dataclass
class LightScene:
    """Inputs of light sub-path tracing: media, nested imaginary regions (outer first; a point
    takes the medium of the last region containing it, else global_medium), point lights and
    the upbp_light_trace_params fields."""
    media: np.ndarray
    regions: np.ndarray  # REGION_DTYPE
    lights: np.ndarray  # POINT_LIGHT_DTYPE
    params: dict


def light_scene(photon_beam_type=SHORT_BEAM, max_path_length=6, beam_path_count=1 << 30, merge_pb2d=1, seed=5,
                n_lights=1, mean_cosine=0.0, bb1d_min_mfp=0.5, media_rows=None) -> LightScene:
    """The box + sphere media (box medium 0 = [0,1]^3, sphere medium 2 of radius 0.3 at the
    centre, vacuum medium 1 outside) lit by point lights inside the box."""
    rows = media_rows or _medium_rows_box_sphere()
    rows = [(a, s, (mean_cosine if i != 1 else 0.0), c) for i, (a, s, g, c) in enumerate(rows)]
    media = make_media(rows)
    c, r, sid = SPHERE
    regions = np.zeros(2, REGION_DTYPE)
    regions[0] = (REGION_BOX, 0, (0, 0, 0), (1, 1, 1))
    regions[1] = (REGION_SPHERE, sid, c, (r, 0, 0))
    lights = np.zeros(n_lights, POINT_LIGHT_DTYPE)
    for i in range(n_lights):
        lights[i] = ((0.5 + 0.3 * np.cos(2.1 * i), 0.5 + 0.3 * np.sin(2.1 * i), 0.1 + 0.05 * i), (1.0, 0.8, 0.6))
    params = dict(seed=seed, photon_beam_type=photon_beam_type, max_path_length=max_path_length, global_medium=1,
                  beam_path_count=beam_path_count, merge_pb2d=merge_pb2d, surf_factor=0.7, pp3d_factor=0.9,
                  pb2d_factor=0.6, bb1d_factor=1.3, bb1d_min_mfp=bb1d_min_mfp)
    return LightScene(media, regions, lights, params)


def light_path_capacity(scene: LightScene, path_count: int):
    """Upper bounds of the vertex and beam counts: max_path_length vertices and
    (max_path_length - 1) rays of at most 2 * regions + 1 pieces per path."""
    L = int(scene.params["max_path_length"])
    if L < 2:
        return 0, 0
    return path_count * L, path_count * (L - 1) * (2 * len(scene.regions) + 1)


def light_path_arrays(vertex_capacity: int, beam_capacity: int, path_count: int, zeros=np.zeros) -> dict:
    """Output arrays of upbp_trace_light_paths (numpy by default)."""
    return {"vertices": zeros(max(vertex_capacity, 1), LIGHT_VERTEX_DTYPE),
            "positions": zeros((max(vertex_capacity, 1), 3), np.float32),
            "path_ends": zeros(max(path_count, 1), np.int32),
            "beam_origin": zeros((max(beam_capacity, 1), 3), np.float32),
            "beam_direction": zeros((max(beam_capacity, 1), 3), np.float32),
            "beam_length": zeros(max(beam_capacity, 1), np.float32),
            "beam_throughput": zeros((max(beam_capacity, 1), 3), np.float32),
            "beam_flags": zeros(max(beam_capacity, 1), np.uint32),
            "beam_medium": zeros(max(beam_capacity, 1), np.int32),
            "beam_mis": zeros(max(beam_capacity, 1), BEAM_MIS_DTYPE)}


def trim_light_paths(arrays: dict, n_vertices: int, n_beams: int, path_count: int) -> dict:
    out = {k: v[:n_vertices] for k, v in arrays.items() if k in ("vertices", "positions")}
    out.update({k: v[:n_beams] for k, v in arrays.items() if k.startswith("beam_")})
    out["path_ends"] = arrays["path_ends"][:path_count]
    return out


def light_paths_to_inputs(lp: dict, media: np.ndarray, beam_radius=0.003, photon_radius=0.005):
    """The estimators' inputs from traced light paths, as the reference's iteration builds
    them (UPBP.hxx:676-730): photons = the in-medium light vertices (position, dir_fix as the
    incoming direction, throughput as power), beams with their MIS records, and the light
    vertices + path ends of the MIS walks. Returns (Photons, in_medium u8, Beams, beam_mis,
    light_vertices, path_ends)."""
    lv = lp["vertices"]
    ph = Photons(np.ascontiguousarray(lp["positions"], np.float32), np.ascontiguousarray(lv["dir_fix"], np.float32),
                 np.ascontiguousarray(lv["throughput"], np.float32), None, photon_radius)
    in_medium = ((lv["flags"] & LV_IN_MEDIUM) != 0).astype(np.uint8)
    bm = Beams(np.ascontiguousarray(lp["beam_origin"]), np.ascontiguousarray(lp["beam_direction"]),
               np.ascontiguousarray(lp["beam_length"]), np.ascontiguousarray(lp["beam_throughput"]),
               np.ascontiguousarray(lp["beam_flags"]), np.ascontiguousarray(lp["beam_medium"]), beam_radius)
    return ph, in_medium, bm, np.ascontiguousarray(lp["beam_mis"]), lv, np.ascontiguousarray(lp["path_ends"])
