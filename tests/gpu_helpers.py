"""Shared helpers of the GPU parity tests (compare the HIP library with the CPU oracle)."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po
from smallupbp_amd import native

# per-query relative error bound against the FP32 oracle on the same inputs
REL_TOL = 1e-4


def rel_err(gpu: np.ndarray, ref: np.ndarray, floor: float = 0.0) -> np.ndarray:
    """Per-query max-channel relative error |gpu - ref| / max(|ref|, floor)."""
    den = np.maximum(np.abs(ref).max(axis=1), floor)
    num = np.abs(gpu.astype(np.float64) - ref.astype(np.float64)).max(axis=1)
    out = np.zeros_like(num)
    nz = den > 0
    out[nz] = num[nz] / den[nz]
    out[~nz] = num[~nz]  # both must be exactly zero where the reference is zero
    return out


def gpu_context(media) -> native.Context:
    ctx = native.Context(0)
    ctx.set_media(media)
    return ctx


def print_case(case: str, **kw):
    print("[case] %s %s" % (case, " ".join("%s=%s" % (k, v) for k, v in kw.items())), flush=True)


def oracle_photons(w, precision=32, accel=1, signatures=True, queries=None):
    oph = po.OraclePhotons(w.photons)
    return po.bre_eval(oph, w.media, queries or w.queries, accel=accel, precision=precision,
                       signatures=signatures)


def oracle_beams(w, precision=32, accel=0, signatures=True, queries=None, grid_res=256):
    obm = po.OracleBeams(w.beams, grid_res=grid_res if accel == 1 else 0)
    return po.beams_eval(obm, w.media, queries or w.queries, accel=accel, precision=precision,
                         signatures=signatures)


def hash32(i):
    """The signature hash of an item index (lowbias32 of i + 1), as the kernels and the oracle."""
    x = (np.asarray(i, np.uint64) + 1) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


def beam_hits_f32(queries, j, beams):
    """The reference beam predicate (PhotonBeam::testIntersectionBeamBeam + the radius test,
    PhotonBeam.hxx:238-334) for query j against every beam, in FP32 with the reference's
    operation order (numpy float32 is IEEE single, no contraction). Returns, per hit, the beam
    index and tw, the foot of the query's closest point on the beam line (the parameter the GPU
    pieces own pairs by)."""
    f = np.float32
    o1 = queries.origin[j].astype(f)
    d1 = queries.direction[j].astype(f)
    O2 = beams.origin.astype(f)
    D2 = beams.direction.astype(f)
    L = beams.length.astype(f)
    r = f(beams.radius)

    def dot(ax, ay, az, bx, by, bz):
        return (ax * bx + ay * by) + az * bz

    cx = d1[1] * D2[:, 2] - d1[2] * D2[:, 1]
    cy = d1[2] * D2[:, 0] - d1[0] * D2[:, 2]
    cz = d1[0] * D2[:, 1] - d1[1] * D2[:, 0]
    s2 = dot(cx, cy, cz, cx, cy, cz)
    ad = dot(O2[:, 0] - o1[0], O2[:, 1] - o1[1], O2[:, 2] - o1[2], cx, cy, cz)
    ok = ~(ad * ad >= (r * r) * s2)
    q = dot(d1[0], d1[1], d1[2], D2[:, 0], D2[:, 1], D2[:, 2])
    q2m1 = q * q - f(1)
    ok &= ~((q2m1 < f(1e-5)) & (q2m1 > f(-1e-5)))
    d1o1 = dot(d1[0], d1[1], d1[2], o1[0], o1[1], o1[2])
    d1o2 = dot(d1[0], d1[1], d1[2], O2[:, 0], O2[:, 1], O2[:, 2])
    d2o2 = dot(D2[:, 0], D2[:, 1], D2[:, 2], O2[:, 0], O2[:, 1], O2[:, 2])
    d2oo = dot(D2[:, 0], D2[:, 1], D2[:, 2], o1[0], o1[1], o1[2]) - d2o2
    with np.errstate(all="ignore"):
        t1 = ((d1o1 - d1o2) - q * d2oo) / q2m1
        t2 = ((t1 + d1o1) - d1o2) / q
    tw = d2oo + t1 * q
    hits = []
    for sg in queries.segments[queries.seg_offsets[j]:queries.seg_offsets[j + 1]]:
        m = ok & (t1 > sg["dist_min"]) & (t1 < sg["dist_max"]) & (t2 > 0) & (t2 < L) & ~np.isnan(t2)
        m &= beams.medium_id == sg["medium_id"]
        sel = np.nonzero(m)[0]
        with np.errstate(all="ignore"):
            dist = np.abs(ad[sel]) / np.sqrt(s2[sel])
        sel = sel[~(r < dist)]
        hits += [(int(b), float(tw[b]), float(L[b]), float(q[b])) for b in sel]
    return hits


def check_beam_ties(queries, beams, sig_gpu, sig_ref):
    """Hit-set differences between the GPU and the oracle are allowed only as the documented
    end ties: the GPU misses (never adds) pairs that the reference accepts although the foot of
    the closest point lies beyond an end of the beam (tw outside [0, length]); the reference's
    t2 = n / q then lands inside (0, length) only through its rounding error, about 1e-7 / |q|
    of the scene scale, for nearly perpendicular lines. Each differing query is re-evaluated
    with the reference predicate in FP32 (checked against the oracle's signature) and the GPU
    signature must equal the oracle's minus a subset of those end pairs. Returns the number of
    tied pairs."""
    import itertools

    bad = np.nonzero(np.any(sig_gpu != sig_ref, axis=1))[0]
    ties = 0
    for j in bad:
        hits = beam_hits_f32(queries, j, beams)
        hs = hash32(np.array([h[0] for h in hits], np.int64)) if hits else np.zeros(0, np.uint32)
        assert len(hits) == int(sig_ref[j, 0]) and int(hs.astype(np.uint64).sum() % (1 << 32)) == int(sig_ref[j, 1]), \
            "query %d: FP32 restatement disagrees with the oracle" % j
        ends = [k for k, h in enumerate(hits) if not (0.0 <= h[1] <= h[2])]
        ok = False
        for n in range(1, len(ends) + 1):
            for sub in itertools.combinations(ends, n):
                keep = [k for k in range(len(hits)) if k not in sub]
                if len(keep) == int(sig_gpu[j, 0]) and \
                        int(hs[keep].astype(np.uint64).sum() % (1 << 32)) == int(sig_gpu[j, 1]):
                    ok = True
                    ties += n
                    break
            if ok:
                break
        assert ok, "query %d: GPU %s vs oracle %s is not an end tie (end pairs %s)" % (
            j, sig_gpu[j], sig_ref[j], [hits[k] for k in ends])
    return ties


def beam_mismatch_report(w, queries, sig_gpu, sig_ref, piece_factor=8.0, max_pieces=64, limit=6):
    """Names the beams behind differing beam hit signatures: for each differing query, the beam
    whose hash accounts for the difference (one beam missing or extra), with its closest-approach
    parameters, the radius margin and the piece holding t2. Diagnostics only."""
    bad = np.nonzero(np.any(sig_gpu != sig_ref, axis=1))[0]
    lines = ["%d of %d queries differ" % (len(bad), len(sig_gpu))]
    bm = w.beams
    O2 = bm.origin.astype(np.float64)
    D2 = bm.direction.astype(np.float64)
    H = hash32(np.arange(bm.count))
    for j in bad[:limit]:
        gn, gh = int(sig_gpu[j, 0]), int(sig_gpu[j, 1])
        rn, rh = int(sig_ref[j, 0]), int(sig_ref[j, 1])
        dh = (gh - rh) % (1 << 32)
        line = "query %d: gpu %d hits, oracle %d" % (j, gn, rn)
        cand = np.nonzero(H == (dh if gn > rn else (-dh) % (1 << 32)))[0] if abs(gn - rn) == 1 else []
        for b in cand[:2]:
            o1 = queries.origin[j].astype(np.float64)
            d1 = queries.direction[j].astype(np.float64)
            c = np.cross(d1, D2[b])
            s2 = float(c @ c)
            ad = float((O2[b] - o1) @ c)
            qd = float(D2[b] @ d1)
            w0 = o1 - O2[b]
            t1 = (qd * float(w0 @ D2[b]) - float(w0 @ d1)) / (1 - qd * qd)
            t2 = (float(w0 @ D2[b]) - qd * float(w0 @ d1)) / (1 - qd * qd)
            ln = float(bm.length[b])
            segs = queries.segments[queries.seg_offsets[j]:queries.seg_offsets[j + 1]]
            inseg = [(float(s["dist_min"]), float(s["dist_max"]), int(s["medium_id"])) for s in segs
                     if s["dist_min"] < t1 < s["dist_max"]]
            npc = int(min(max(np.ceil(ln / (piece_factor * bm.radius)), 1), max_pieces))
            k = t2 / ln * npc
            line += ("; beam %d (%s): t1 %.7g in %s, t2 %.7g of length %.7g (piece %.4f of %d), "
                     "dist/r %.8f, medium %d" % (b, "missing" if rn > gn else "extra", t1, inseg, t2, ln, k, npc,
                                                 abs(ad) / np.sqrt(s2) / bm.radius, int(bm.medium_id[b])))
        lines.append(line)
    return "\n".join(lines)
