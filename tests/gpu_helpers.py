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
