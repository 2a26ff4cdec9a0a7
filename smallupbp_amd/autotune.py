"""Autotuner of the estimator kernels.

Searches the parameters that the GPU path exposes and keeps the fastest per
workload-size bucket:
  build-time: LBVH leaf size (items per leaf tile, 4..64) for photons and for
              beam pieces, and the beam piece length (in units of the beam radius);
  launch-time: waves per workgroup (1, 2, 4) and the query order (Morton-sorted
              bundles or the caller's order); with both estimators, the photon kernel's
              waves per workgroup once more on the joint (concurrent) step.
Times come from the library's HIP events around each traversal kernel (median of
`reps` runs). Results are cached in a JSON file keyed by
(log2 photons, log2 beams, log2 queries) so later runs of similar size reuse them.
"""
from __future__ import annotations

import itertools
import json
import math
import os
from typing import Dict, Optional

import numpy as np

from . import _abi, native, scenes

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

# Leaves up to 64 items (the library's limit): when a wave's rays share their origin, a leaf's
# items are checked against the wave's frustum in one pass, so large leaves can pay off.
PHOTON_LEAF = (4, 8, 16, 32, 64)
BEAM_LEAF = (4, 8, 16, 32, 64)
PIECE_FACTOR = (4.0, 8.0, 16.0, 32.0)
WAVES = (1, 2, 4)
ORDERS = (_abi.ORDER_SORTED, _abi.ORDER_AS_GIVEN)


def bucket(n_photons: int, n_beams: int, n_queries: int) -> str:
    def b(n):
        return 0 if n <= 0 else int(round(math.log2(n)))

    return "%d/%d/%d" % (b(n_photons), b(n_beams), b(n_queries))


def load_cache(path: str = DEFAULT_CACHE) -> Dict[str, dict]:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(n_photons: int, n_beams: int, n_queries: int, path: str = DEFAULT_CACHE) -> Optional[dict]:
    return load_cache(path).get(bucket(n_photons, n_beams, n_queries))


def tuned_params(n_photons: int, n_beams: int, n_queries: int, path: str = DEFAULT_CACHE) -> dict:
    """Build and launch parameters of the cached best configuration for this size bucket
    (0 / "sorted" = library default where the bucket has no entry)."""
    t = lookup(n_photons, n_beams, n_queries, path) or {}
    ph = t.get("best", {}).get("photons") or {}
    bm = t.get("best", {}).get("beams") or {}
    return {"photon_leaf_size": int(ph.get("leaf_size", 0)), "beam_leaf_size": int(bm.get("leaf_size", 0)),
            "piece_factor": float(bm.get("piece_factor", 0.0)),
            "photon_waves_per_block": int(ph.get("waves_per_block", 0)),
            "beam_waves_per_block": int(bm.get("waves_per_block", 0)),
            "order": _abi.ORDER_AS_GIVEN if bm.get("order") == _abi.ORDER_AS_GIVEN else _abi.ORDER_SORTED,
            "cached": bool(t)}


def _median_kernel_ms(ctx, bre, beams, qv, reps, which):
    ts = []
    for _ in range(reps):
        st = {}
        native.eval_batch(ctx, bre if which == "photons" else None, beams if which == "beams" else None, qv,
                          stats=st)
        ts.append(st["kernel_ms_photons"] + st["kernel_ms_beams"] + st["kernel_ms_other"])
    return float(np.median(ts))


def _median_step_ms(ctx, bre, beams, qv, reps):
    """Median wall time of a whole eval call (both kernels, synchronised) over `reps` runs."""
    import time

    import torch

    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        native.eval_batch(ctx, bre, beams, qv)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def tune(w: scenes.Workload, reps: int = 3, path: Optional[str] = DEFAULT_CACHE, log=print) -> dict:
    """Sweep the parameter space for workload `w`; returns (and caches) the best configuration."""
    ctx = native.Context(0)
    ctx.set_media(w.media)
    best = {"photons": None, "beams": None}
    results = {"photons": [], "beams": []}
    qviews = {o: native.QueryBatchView.to_device(w.queries, order=o) for o in ORDERS}
    if w.photons is not None:
        for leaf in PHOTON_LEAF:
            bre = native.EmbreeBre(ctx)
            bre.build(w.photons, leaf_size=leaf)
            for wpb, order in itertools.product(WAVES, ORDERS):
                ctx.set_tuning(wpb)
                native.eval_batch(ctx, bre, None, qviews[order])  # warm
                ms = _median_kernel_ms(ctx, bre, None, qviews[order], reps, "photons")
                cfg = {"leaf_size": leaf, "waves_per_block": wpb, "order": order, "ms": ms}
                results["photons"].append(cfg)
                log("photons %s" % cfg)
                if best["photons"] is None or ms < best["photons"]["ms"]:
                    best["photons"] = cfg
            bre.destroy()
    if w.beams is not None:
        for leaf, pf in itertools.product(BEAM_LEAF, PIECE_FACTOR):
            bm = native.PhotonBeamsEvaluator(ctx)
            bm.build(w.beams, leaf_size=leaf, piece_factor=pf)
            for wpb, order in itertools.product(WAVES, ORDERS):
                ctx.set_tuning(wpb)
                native.eval_batch(ctx, None, bm, qviews[order])
                ms = _median_kernel_ms(ctx, None, bm, qviews[order], reps, "beams")
                cfg = {"leaf_size": leaf, "piece_factor": pf, "pieces": bm.pieces, "waves_per_block": wpb,
                       "order": order, "ms": ms}
                results["beams"].append(cfg)
                log("beams %s" % cfg)
                if best["beams"] is None or ms < best["beams"]["ms"]:
                    best["beams"] = cfg
            bm.destroy()
    if best["photons"] is not None and best["beams"] is not None and best["photons"]["order"] == best["beams"]["order"]:
        # One eval call runs the two kernels concurrently on two streams, and how their blocks
        # share the CUs depends on both launch shapes: a photon shape that ties with others alone
        # can lose 10 % of the joint step. Re-choose the photon waves per block on the joint step.
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons, leaf_size=best["photons"]["leaf_size"])
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams, leaf_size=best["beams"]["leaf_size"], piece_factor=best["beams"]["piece_factor"])
        qv = qviews[best["beams"]["order"]]
        joint = []
        for wpb in WAVES:
            ctx.set_tuning(best["beams"]["waves_per_block"], 0, wpb)
            native.eval_batch(ctx, bre, bm, qv)  # warm
            ms = _median_step_ms(ctx, bre, bm, qv, reps)
            joint.append({"photon_waves_per_block": wpb, "ms": ms})
            log("joint step %s" % joint[-1])
        pick = min(joint, key=lambda j: j["ms"])
        best["photons"] = dict(best["photons"], waves_per_block=pick["photon_waves_per_block"], joint_ms=pick["ms"])
        results["joint"] = joint
        bre.destroy()
        bm.destroy()
    key = bucket(w.photons.count if w.photons is not None else 0, w.beams.count if w.beams is not None else 0,
                 w.queries.count)
    entry = {"workload": w.name, "best": best, "all": results}
    if path:
        cache = load_cache(path)
        cache[key] = entry
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    return entry


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="autotune the estimator kernels on a named workload")
    ap.add_argument("--workload", default="bre_beams_1080p",
                    choices=["bre_beams_1080p", "plain_bre_256", "beams_stress", "adjoint_512"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    a = ap.parse_args()
    w = {"bre_beams_1080p": scenes.bre_beams, "plain_bre_256": scenes.plain_bre,
         "beams_stress": scenes.beams_stress, "adjoint_512": scenes.adjoint_scene}[a.workload]()
    e = tune(w, a.reps, a.cache)
    print(json.dumps(e["best"]))
