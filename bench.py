#!/usr/bin/env python3
"""Benchmark of the volumetric merging estimators on one MI355X.

Workload (default): box + nested-sphere media, 1920x1080 LONG query beams,
1,000,000 photon points (lognormal radii) + 200,000 SHORT photon beams
(radius 0.003), FP32, synthetic seeded data. One step = one evaluation of both
estimators (PB2D photon points and BB1D photon beams) over all queries, with
inputs already resident in HBM (query regrouping by Morton key included in
the step). Structures are built once before timing (the reference also builds
once per iteration and evaluates every camera ray).

Prints ONE JSON line: metric, value (Mqueries/s), roofline of the dominant
kernel (algorithmic bytes / average HIP-event kernel time vs 8 TB/s), the CPU
oracle timed on a bounded sample on the host cores, the forward+adjoint vs
forward cost ratio on the 512x512 differentiable workload, the 64-instance
batched speedup, the MIS-weighted (UPBP) rate at 1920x1080, the light sub-path
tracing rate, the beam estimator under each per-cell beam reduction type, and the
PP3D and SURF vertex-merging rates.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes_queries(q) -> int:
    """Per-query algorithmic bytes: 24 B ray + 12 B per segment (tmin, tmax, medium) + 12 B RGB out."""
    nseg = int(q.seg_offsets[-1])
    return 24 * q.count + 12 * nseg + 12 * q.count


def host_cpu_info() -> dict:
    """The host CPUs this run may use: the process's affinity mask, the machine's logical CPU
    count, and the CPU share the job is given (OMP_NUM_THREADS, which the GPU host sets to the
    cores it allots per GPU)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    share_n = int(share) if share.isdigit() and int(share) > 0 else affinity
    return {"affinity": affinity, "os_cpu_count": os.cpu_count(), "job_share": share_n,
            "threads": max(1, min(affinity, share_n)), "model": cpu_name()}


def host_cores() -> int:
    """Oracle threads: every CPU of the affinity mask within the job's CPU share."""
    return host_cpu_info()["threads"]


def cpu_name() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(w, budget_s: float, threads: int):
    """Oracle (clean-room CPU restatement of the reference estimator) on a random query sample:
    photons through a binned-SAH BVH, beams through the reference uniform grid (256 cells on the
    longest axis, 3D-DDA). Returns Mqueries/s and what the sample was."""
    from oracle import pyoracle as po

    t = time.time()
    oph = po.OraclePhotons(w.photons)
    obm = po.OracleBeams(w.beams, grid_res=256)
    build_s = time.time() - t
    rng = np.random.default_rng(12345)
    n_probe = min(w.queries.count, 20000)
    probe = np.sort(rng.choice(w.queries.count, n_probe, replace=False))
    qp = w.queries.subset(probe)
    po.bre_eval(oph, w.media, qp, accel=1, threads=threads)  # warm the structures
    t = time.time()
    po.bre_eval(oph, w.media, qp, accel=1, threads=threads)
    po.beams_eval(obm, w.media, qp, accel=1, threads=threads)
    per_q = (time.time() - t) / n_probe
    n = int(min(w.queries.count, max(2000, budget_s / max(per_q, 1e-9))))
    idx = np.sort(rng.choice(w.queries.count, n, replace=False))
    qs = w.queries.subset(idx)
    t = time.time()
    po.bre_eval(oph, w.media, qs, accel=1, threads=threads)
    po.beams_eval(obm, w.media, qs, accel=1, threads=threads)
    el = time.time() - t
    return n / el / 1e6, n, el, build_s


def time_eval(ctx, bre, beams, qv, out_p, out_b, steps, warmup):
    import torch

    from smallupbp_amd import native

    for _ in range(warmup):
        native.eval_batch(ctx, bre, beams, qv, out_p, out_b)
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = {}
        native.eval_batch(ctx, bre, beams, qv, out_p, out_b, stats=st)
        per.append(st)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return el, per


def adjoint_ratio(steps: int):
    """Forward vs forward + adjoint on the differentiable workload: 512x512 queries x 500k photons
    + 100k beams in the box + sphere media, loss = sum_q |L_q - T_q|^2 against a target image T
    rendered at sigma_s x 1.1. The forward evaluates both estimators; the adjoint step also forms
    the upstream gradient 2 (L - T) and runs the adjoint kernels for the 18 medium coefficients
    and the 500k x 3 photon powers (and the beam powers)."""
    import torch

    from smallupbp_amd import native, scenes

    w = scenes.adjoint_scene()
    tgt = scenes.with_media(w, scenes.scale_scattering(w.media, 1.1))
    ctx = native.Context(0)
    ctx.set_media(tgt.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    beams = native.PhotonBeamsEvaluator(ctx)
    beams.build(w.beams)
    tp, tb = native.eval_batch(ctx, bre, beams, native.QueryBatchView.to_device(tgt.queries))
    target = (tp + tb).contiguous()
    ctx.set_media(w.media)
    qv = native.QueryBatchView.to_device(w.queries)
    dev = "cuda"
    op = torch.empty((w.queries.count, 3), device=dev)
    ob = torch.empty((w.queries.count, 3), device=dev)
    gs = torch.zeros((len(w.media), 3), device=dev)
    ga = torch.zeros((len(w.media), 3), device=dev)
    gpp = torch.zeros((w.photons.count, 3), device=dev)
    gbp = torch.zeros((w.beams.count, 3), device=dev)

    def fwd_adj():
        native.eval_batch(ctx, bre, beams, qv, op, ob)
        g = 2.0 * (op + ob - target)
        native.adjoint(ctx, bre, beams, qv, g, g, gs, ga, gpp, gbp)

    for _ in range(2):
        native.eval_batch(ctx, bre, beams, qv, op, ob)
        fwd_adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        native.eval_batch(ctx, bre, beams, qv, op, ob)
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        fwd_adj()
    torch.cuda.synchronize()
    t_fa = (time.perf_counter() - t0) / steps
    return t_fa / t_f, t_f * 1e3, t_fa * 1e3


def stress_measure(steps: int, cpu_budget: float, threads: int):
    """Photon-beam-only stress: 1,000,000 query beams from random camera rays x 2,000,000 photon
    beams in one medium (sigma_s 1, sigma_a 0.1; half SHORT with Exp(mean 0.05) lengths, half
    LONG to the box boundary), radius 0.003. Inputs resident in HBM; one step = one evaluation of
    the beam estimator over all queries (query regrouping included). Roofline of the beam kernel
    from its HIP-event time; the oracle through the reference's uniform grid on a query sample is
    the CPU comparison."""
    import torch

    from smallupbp_amd import autotune, native, scenes

    t = time.time()
    w = scenes.beams_stress()
    log("stress workload %s generated in %.1fs" % (w.name, time.time() - t))
    p = autotune.tuned_params(0, w.beams.count, w.queries.count)
    ctx = native.Context(0)
    ctx.set_media(w.media)
    if p["beam_waves_per_block"]:
        ctx.set_tuning(p["beam_waves_per_block"])
    bm = native.PhotonBeamsEvaluator(ctx)
    torch.cuda.synchronize()
    t = time.perf_counter()
    bm.build(w.beams, leaf_size=p["beam_leaf_size"], piece_factor=p["piece_factor"])
    t_build = time.perf_counter() - t
    qv = native.QueryBatchView.to_device(w.queries, order=p["order"])
    Q = w.queries.count
    out = torch.empty((Q, 3), device="cuda")
    el, per = time_eval(ctx, None, bm, qv, None, out, steps, 2)
    ms_bm = float(np.mean([x["kernel_ms_beams"] for x in per]))
    H = int(per[-1]["beam_hits"])
    tests = int(per[-1]["beam_candidates"])
    dom_bytes = alg_bytes_queries(w.queries) + 48 * H
    achieved = dom_bytes / (ms_bm * 1e-3) / 1e9
    traffic = None
    try:
        traffic = json.load(open(os.path.join(ROOT, "profiles", "traffic.json"))).get("k_eval<beams> stress")
    except (OSError, ValueError):
        traffic = None
    rec = {"workload": "beams_stress (1M query beams x 2M photon beams, 50 % SHORT / 50 % LONG, r = 0.003)",
           "mqueries_per_s": round(Q / (el / steps) / 1e6, 3), "ms_per_step": round(el / steps * 1e3, 4),
           "kernel_ms": round(ms_bm, 4), "beam_hits": H, "hits_per_query": round(H / Q, 1),
           "beam_tests": tests, "tests_per_hit": round(tests / max(H, 1), 2), "pieces": bm.pieces,
           "build_ms": round(t_build * 1e3, 2),
           "tuning": {k: v for k, v in p.items() if k != "cached"} if p["cached"] else "library defaults",
           "roofline": {"bound": "hbm", "kernel": "k_eval<beams>", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                        "traffic": traffic, "alg_bytes_per_launch": dom_bytes}}
    if cpu_budget > 0:
        from oracle import pyoracle as po

        t = time.time()
        obm = po.OracleBeams(w.beams, grid_res=256)
        build_s = time.time() - t
        rng = np.random.default_rng(31)
        probe = np.sort(rng.choice(Q, 500, replace=False))
        t = time.time()
        po.beams_eval(obm, w.media, w.queries.subset(probe), accel=1, threads=threads)
        per_q = (time.time() - t) / len(probe)
        n = int(min(Q, max(500, cpu_budget / max(per_q, 1e-9))))
        idx = np.sort(rng.choice(Q, n, replace=False))
        t = time.time()
        po.beams_eval(obm, w.media, w.queries.subset(idx), accel=1, threads=threads)
        el_cpu = time.time() - t
        rec["cpu_baseline"] = {"value": round(n / el_cpu / 1e6, 6), "unit": "Mqueries/s", "cores": threads,
                               "kind": "port", "sample": "%d random queries of the same workload through the oracle's "
                               "restatement of the reference uniform grid (256 cells, 3D-DDA); %.1f s on %d threads, "
                               "grid built in %.1f s" % (n, el_cpu, threads, build_s)}
        rec["gpu_vs_cpu"] = round(rec["mqueries_per_s"] / rec["cpu_baseline"]["value"], 1)
    bm.destroy()
    return rec


def iteration_measure(steps: int, ctx, qv, leaf_ph: int, leaf_bm: int, piece_factor: float, paths: int = 1_000_000,
                      beam_paths: int = 125_000):
    """One renderer iteration on the GPU with every array device-resident: trace 1,000,000 light
    paths in the box + sphere media (upbp_trace_light_paths: light vertices, and SHORT beams
    from the first 125,000 paths, the reference's `-pbc 125000`, about 200k beams as in the
    1080p workload), build the photon structure from the in-medium light vertices (r = 0.005)
    and the beam structure (r = 0.003) from device memory, then evaluate both estimators for
    the 1920x1080 query beams. The reference rebuilds its structures every iteration
    (SmallUPBP/src/Renderers/VolLightTracer.hxx:303-410), so this is the per-iteration cost."""
    import ctypes as C

    import torch

    from smallupbp_amd import _abi, native, scenes

    sc = scenes.light_scene(media_rows=LIGHT_TRACE_MEDIA, max_path_length=10, bb1d_min_mfp=0.1, seed=12,
                            beam_path_count=beam_paths)
    tctx = native.Context(0)
    tctx.set_media(sc.media)
    cap_v, cap_b = scenes.light_path_capacity(sc, paths)

    def zeros(n, dtype):
        return torch.zeros(max(int(n), 1) * np.dtype(dtype).itemsize, dtype=torch.uint8, device="cuda")

    proto = scenes.light_path_arrays(1, 1, 1)
    arr = {k: zeros((cap_v if k in ("vertices", "positions") else paths if k == "path_ends" else cap_b)
                    * int(np.prod(v.shape[1:])), v.dtype) for k, v in proto.items()}
    lp = _abi.light_paths(arr, cap_v, cap_b)
    regions = np.ascontiguousarray(sc.regions)
    lights = np.ascontiguousarray(sc.lights)
    par = _abi.LightTraceParams()
    for k, v in sc.params.items():
        setattr(par, k, v)
    lv_words = scenes.LIGHT_VERTEX_DTYPE.itemsize // 4
    off_tp = scenes.LIGHT_VERTEX_DTYPE.fields["throughput"][1] // 4
    off_dir = scenes.LIGHT_VERTEX_DTYPE.fields["dir_fix"][1] // 4
    off_fl = scenes.LIGHT_VERTEX_DTYPE.fields["flags"][1] // 4
    out_p = torch.empty((qv.count, 3), device="cuda")
    out_b = torch.empty((qv.count, 3), device="cuda")
    ctx.set_media(sc.media)

    def one(times=None):
        tctx.check(tctx.lib.upbp_trace_light_paths(tctx.h, regions.ctypes.data, len(regions), lights.ctypes.data,
                                                   len(lights), C.byref(par), paths, C.byref(lp), _abi.MEM_DEVICE))
        t1 = time.perf_counter()
        nv, nb = int(lp.vertex_count), int(lp.beam_count)
        # photon inputs from the light vertices: strided views of the device records
        words = arr["vertices"][: nv * lv_words * 4].view(torch.float32).view(nv, lv_words)
        power = words[:, off_tp:off_tp + 3].contiguous()
        inc = words[:, off_dir:off_dir + 3].contiguous()
        in_med = ((words[:, off_fl].view(torch.int32) & scenes.LV_IN_MEDIUM) != 0).to(torch.uint8)
        pos = arr["positions"][: nv * 12].view(torch.float32).view(nv, 3)
        bre = native.EmbreeBre(ctx)
        bre.build(scenes.Photons(pos, inc, power, None, 0.005), in_medium=in_med, leaf_size=leaf_ph)
        t2 = time.perf_counter()
        bmv = {k: arr[k][: nb * int(np.prod(proto[k].shape[1:])) * proto[k].dtype.itemsize].view(
            {np.dtype(np.float32): torch.float32, np.dtype(np.uint32): torch.int32,
             np.dtype(np.int32): torch.int32}[proto[k].dtype]) for k in
            ("beam_origin", "beam_direction", "beam_length", "beam_throughput", "beam_flags", "beam_medium")}
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(scenes.Beams(bmv["beam_origin"].view(nb, 3), bmv["beam_direction"].view(nb, 3),
                              bmv["beam_length"], bmv["beam_throughput"].view(nb, 3), bmv["beam_flags"],
                              bmv["beam_medium"], 0.003), leaf_size=leaf_bm, piece_factor=piece_factor)
        t3 = time.perf_counter()
        st = {}
        native.eval_batch(ctx, bre, bm, qv, out_p, out_b, stats=st)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        if times is not None:
            times.append((t1, t2, t3, t4))
            hits.append((int(st["photon_hits"]), int(st["beam_hits"])))
        kept = bre.num_photons
        bre.destroy()
        bm.destroy()
        return nv, nb, kept

    hits = []
    for _ in range(2):
        one()
    torch.cuda.synchronize()
    tms = []
    starts = []
    for _ in range(steps):
        torch.cuda.synchronize()
        starts.append(time.perf_counter())
        nv, nb, kept = one(tms)
    trace = [t[0] - s for t, s in zip(tms, starts)]
    build_ph = [t[1] - t[0] for t in tms]
    build_bm = [t[2] - t[1] for t in tms]
    ev = [t[3] - t[2] for t in tms]
    total = [t[3] - s for t, s in zip(tms, starts)]
    ms = lambda xs: round(float(np.mean(xs)) * 1e3, 3)
    return {"workload": "iteration (trace %d light paths -> %d light vertices (%d in media, the photons) + %d SHORT "
                        "beams from the first %d paths -> photon and beam builds from device memory -> eval of "
                        "1920x1080 queries)" % (paths, nv, kept, nb, beam_paths),
            "iteration_ms": ms(total), "trace_ms": ms(trace), "photon_build_ms": ms(build_ph),
            "beam_build_ms": ms(build_bm), "eval_ms": ms(ev), "photon_hits": hits[-1][0], "beam_hits": hits[-1][1]}


def batched_ratio(scale: float, steps: int, n_instances: int = 64):
    """64 independent uneven instances (photons lognormal 10k-2M, beams N/5, queries U[16k, 2M]):
    per-instance upbp_eval launches back to back vs one persistent launch (eager and hipGraph).
    Both evaluate the same structures and queries; structures are built beforehand."""
    import torch

    from smallupbp_amd import native, scenes

    t = time.time()
    ws = scenes.batched_instances(n_instances=n_instances, seed=4, scale=scale, max_queries=2_000_000)
    log("batched workload (%d instances, scale %.3g) generated in %.1fs" % (len(ws), scale, time.time() - t))
    media = np.concatenate([w.media for w in ws])
    ctx = native.Context(0)
    ctx.set_media(media)
    # per-instance contexts hold the instance's own media for the plain launches (media base 0)
    inst, plain = [], []
    for i, w in enumerate(ws):
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons)
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams)
        qv = native.QueryBatchView.to_device(w.queries)
        op = torch.empty((w.queries.count, 3), device="cuda")
        ob = torch.empty((w.queries.count, 3), device="cuda")
        inst.append((bre, bm, qv, 2 * i, op, ob))
        pctx = native.Context(0)
        pctx.set_media(w.media)
        pbre = native.EmbreeBre(pctx)
        pbre.build(w.photons)
        pbm = native.PhotonBeamsEvaluator(pctx)
        pbm.build(w.beams)
        plain.append((pctx, pbre, pbm, qv, op, ob))
    plan = native.BatchPlan(ctx, inst)
    Q = sum(w.queries.count for w in ws)

    def run_plain():
        for pctx, pbre, pbm, qv, op, ob in plain:
            native.eval_batch(pctx, pbre, pbm, qv, op, ob)

    res = {}
    for name, fn in (("plain", run_plain), ("batched", lambda: plan.run(use_graph=False)),
                     ("batched_graph", lambda: plan.run(use_graph=True))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps * 1e3
    return {"workload": "batched_64 (%d instances, scale %.3g, %d queries total)" % (len(ws), scale, Q),
            "plain_ms": round(res["plain"], 3), "batched_ms": round(res["batched"], 3),
            "batched_graph_ms": round(res["batched_graph"], 3),
            "speedup": round(res["plain"] / min(res["batched"], res["batched_graph"]), 3)}


def vm_measure(technique: int, steps: int, cpu_budget: float, threads: int):
    """Vertex merging through the reference's hashed grid with RangeQuery::Process and UPBP MIS
    weights: PP3D (1,000,000 camera vertices in the box + sphere media against the in-medium
    vertices of 1,000,000 light paths) or SURF (1,000,000 camera vertices on three faces of the box
    against the surface vertices of 1,000,000 light paths); radius 0.01, one hash bucket per light
    path. Inputs resident in HBM; one step = one evaluation call (lane sort + grid walk + weights).
    The oracle's restatement of the same hashed grid is timed on a query sample as the CPU
    comparison."""
    import torch

    from smallupbp_amd import native, scenes

    pp3d = technique == scenes.PP3D
    name = "pp3d" if pp3d else "surf"
    t = time.time()
    gen = scenes.pp3d_scene if pp3d else scenes.surf_scene
    sc = gen(n_paths=1_000_000, n_queries=1_000_000, radius=0.01, cell_count=1_000_000, seed=91 if pp3d else 93)
    log("%s workload: %d light vertices, %d camera vertices, generated in %.1fs"
        % (name, len(sc.positions), len(sc.queries), time.time() - t))
    ctx = native.Context(0)
    ctx.set_media(sc.media)
    pos = torch.from_numpy(np.ascontiguousarray(sc.positions)).cuda()
    lvt = torch.from_numpy(np.ascontiguousarray(sc.mis.light_vertices).view(np.uint8)).cuda()
    grid = native.HashGrid(ctx)
    torch.cuda.synchronize()
    t = time.perf_counter()
    grid.build(pos, lvt, sc.radius, sc.cell_count, technique)
    t_build = time.perf_counter() - t
    ms = native.device_mis(sc.mis)
    q = torch.from_numpy(np.ascontiguousarray(sc.queries).view(np.uint8)).cuda()
    Q = len(sc.queries)
    out = torch.empty((Q, 3), device="cuda")
    for _ in range(2):
        grid.eval(q, ms, out=out)
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = {}
        grid.eval(q, ms, out=out, stats=st)
        per.append(st)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_step = el / steps * 1e3
    where = "in-medium" if pp3d else "surface"
    rec = {"workload": "%s_1m (1M camera vertices x %d %s light vertices, r = 0.01, MIS on)"
                       % (name, grid.count, where),
           "mqueries_per_s": round(Q / (el / steps) / 1e6, 3), "ms_per_step": round(ms_step, 4),
           "kernel_ms": round(float(np.mean([p["kernel_ms_photons"] for p in per])), 4),
           "process_calls": int(per[-1]["photon_hits"]), "distance_tests": int(per[-1]["photon_candidates"]),
           "build_ms": round(t_build * 1e3, 2)}
    if cpu_budget > 0:
        from oracle import pyoracle as po

        og = po.OracleHashGrid(sc)
        rng = np.random.default_rng(7)
        probe = np.sort(rng.choice(Q, 20000, replace=False))
        t = time.time()
        po.vm_eval(og, sc, queries=sc.queries[probe], threads=threads)
        per_q = (time.time() - t) / len(probe)
        n = int(min(Q, max(20000, cpu_budget / max(per_q, 1e-9))))
        idx = np.sort(rng.choice(Q, n, replace=False))
        t = time.time()
        po.vm_eval(og, sc, queries=sc.queries[idx], threads=threads)
        el_cpu = time.time() - t
        rec["cpu_baseline"] = {"value": round(n / el_cpu / 1e6, 5), "unit": "Mqueries/s", "cores": threads,
                               "kind": "port", "sample": "%d random camera vertices of the same workload through "
                               "the oracle's restatement of the reference hashed grid; %.1f s on %d threads"
                               % (n, el_cpu, threads)}
        rec["gpu_vs_cpu"] = round(rec["mqueries_per_s"] / rec["cpu_baseline"]["value"], 1)
    grid.destroy()
    return rec


def upbp_measure(steps: int, cpu_budget: float, threads: int):
    """The same estimators with UPBP MIS weights (upbp_eval_upbp) at 1920x1080 LONG query beams:
    400,000 light paths of 1-4 segments (1.0M in-medium light vertices, the photons) and
    200,000 SHORT photon beams leaving random light vertices, photon radius 0.005, beam radius
    0.003; camera sub-paths of 1-3 vertices per query with random MisData. Inputs resident in
    HBM; one step = one UPBP evaluation of both estimators. The plain (unweighted) evaluation
    of the same structures is timed beside it, and the oracle's UPBP evaluation on a query
    sample is the CPU comparison."""
    import torch

    from smallupbp_amd import native, scenes

    t = time.time()
    w, inm, mi = scenes.upbp_scene(1920, 1080, 400_000, 200_000, seed=81, photon_radius=0.005, beam_radius=0.003)
    log("upbp workload: %d light vertices (%d photons), %d beams, %d queries, generated in %.1fs"
        % (len(mi.light_vertices), int(inm.sum()), w.beams.count, w.queries.count, time.time() - t))
    ctx = native.Context(0)
    ctx.set_media(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons, in_medium=inm)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries)
    ms = native.device_mis(mi)
    Q = w.queries.count
    op = torch.empty((Q, 3), device="cuda")
    ob = torch.empty((Q, 3), device="cuda")

    def run(upbp, st=None):
        if upbp:
            native.eval_upbp(ctx, bre, bm, qv, ms, stats=st, out_photons=op, out_beams=ob)
        else:
            native.eval_batch(ctx, bre, bm, qv, op, ob, stats=st)

    res = {}
    for name in ("plain", "upbp"):
        for _ in range(2):
            run(name == "upbp")
        torch.cuda.synchronize()
        per = []
        t0 = time.perf_counter()
        for _ in range(steps):
            st = {}
            run(name == "upbp", st)
            per.append(st)
        torch.cuda.synchronize()
        res[name] = ((time.perf_counter() - t0) / steps, per)
    el, per = res["upbp"]
    rec = {"workload": "upbp_1080p (%d LONG queries x %d photon light vertices + %d SHORT beams, MIS on)"
                       % (Q, int(inm.sum()), w.beams.count),
           "mqueries_per_s": round(Q / el / 1e6, 3), "ms_per_step": round(el * 1e3, 4),
           "kernel_ms": {"photons": round(float(np.mean([p["kernel_ms_photons"] for p in per])), 4),
                         "beams": round(float(np.mean([p["kernel_ms_beams"] for p in per])), 4)},
           "hits": {"photons": int(per[-1]["photon_hits"]), "beams": int(per[-1]["beam_hits"])},
           "plain_ms_per_step": round(res["plain"][0] * 1e3, 4),
           "mis_cost": round(el / res["plain"][0], 3)}
    if cpu_budget > 0:
        from oracle import pyoracle as po

        t = time.time()
        oph = po.OraclePhotons(w.photons, in_medium=inm)
        obm = po.OracleBeams(w.beams, grid_res=256)
        build_s = time.time() - t
        rng = np.random.default_rng(9)

        def cpu(idx):
            qs = w.queries.subset(idx)
            sub = scenes.MisInputs(mi.queries[idx], mi.camera, mi.light_vertices, mi.path_ends, mi.beams, mi.params)
            t = time.time()
            po.upbp_eval(oph, None, w.media, qs, sub, 0, threads=threads, accel=1)
            po.upbp_eval(None, obm, w.media, qs, sub, 1, threads=threads, accel=1)
            return time.time() - t

        probe = np.sort(rng.choice(Q, 2000, replace=False))
        per_q = cpu(probe) / len(probe)
        n = int(min(Q, max(2000, cpu_budget / max(per_q, 1e-9))))
        idx = np.sort(rng.choice(Q, n, replace=False))
        el_cpu = cpu(idx)
        rec["cpu_baseline"] = {"value": round(n / el_cpu / 1e6, 5), "unit": "Mqueries/s", "cores": threads,
                               "kind": "port", "sample": "%d random queries of the same workload through the oracle's "
                               "UPBP evaluation (BVH photons, reference beam grid); %.1f s on %d threads, structures "
                               "built in %.1f s" % (n, el_cpu, threads, build_s)}
        rec["gpu_vs_cpu"] = round(rec["mqueries_per_s"] / rec["cpu_baseline"]["value"], 1)
    return rec


LIGHT_TRACE_MEDIA = [((0.05, 0.1, 0.02), (1.2, 1.0, 0.8), 0.0, 0.9), ((0, 0, 0), (0, 0, 0), 0.0, 1.0),
                     ((0.1, 0.05, 0.2), (3.0, 2.5, 2.0), 0.0, 0.8)]


def light_trace_measure(steps: int, cpu_budget: float, threads: int, paths: int = 1_000_000):
    """Light sub-path tracing (upbp_trace_light_paths): 1,000,000 paths from a point light in
    the box + sphere media (sigma_s about 1 and 2.5), up to 10 vertices, SHORT photon beams.
    Outputs resident in HBM and allocated once; one step = one call (trace, scan, pack). The
    oracle's restatement with the same random numbers, on a path sample, is the CPU
    comparison."""
    import ctypes as C

    import torch

    from smallupbp_amd import _abi, native, scenes

    sc = scenes.light_scene(media_rows=LIGHT_TRACE_MEDIA, max_path_length=10, bb1d_min_mfp=0.1, seed=12)
    ctx = native.Context(0)
    ctx.set_media(sc.media)
    cap_v, cap_b = scenes.light_path_capacity(sc, paths)

    def zeros(shape, dtype):
        n = int(np.prod(shape)) if isinstance(shape, tuple) else int(shape)
        return torch.zeros(max(n, 1) * np.dtype(dtype).itemsize, dtype=torch.uint8, device="cuda")

    arr = {k: zeros((cap_v if k in ("vertices", "positions") else paths if k == "path_ends" else cap_b,)
                    + v.shape[1:], v.dtype) for k, v in scenes.light_path_arrays(1, 1, 1).items()}
    lp = _abi.light_paths(arr, cap_v, cap_b)
    regions = np.ascontiguousarray(sc.regions)
    lights = np.ascontiguousarray(sc.lights)
    par = _abi.LightTraceParams()
    for k, v in sc.params.items():
        setattr(par, k, v)

    def run():
        ctx.check(ctx.lib.upbp_trace_light_paths(ctx.h, regions.ctypes.data, len(regions), lights.ctypes.data,
                                                 len(lights), C.byref(par), paths, C.byref(lp), _abi.MEM_DEVICE))

    for _ in range(2):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    rec = {"workload": "lighttrace_1m (%d light paths, box + sphere media, max path length 10, SHORT beams)" % paths,
           "mpaths_per_s": round(paths / el / 1e6, 3), "ms_per_step": round(el * 1e3, 4),
           "vertices": int(lp.vertex_count), "beams": int(lp.beam_count),
           "mvertices_per_s": round(lp.vertex_count / el / 1e6, 3)}
    if cpu_budget > 0:
        from oracle import pyoracle as po

        t = time.time()
        po.trace_light_paths(sc, 20000, threads=threads)
        per = (time.time() - t) / 20000
        n = int(min(paths, max(20000, cpu_budget / max(per, 1e-9))))
        t = time.time()
        po.trace_light_paths(sc, n, threads=threads)
        el_cpu = time.time() - t
        rec["cpu_baseline"] = {"value": round(n / el_cpu / 1e6, 5), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port", "sample": "the first %d paths of the same workload through the oracle's "
                               "restatement (same random numbers), %.1f s on %d threads" % (n, el_cpu, threads)}
        rec["gpu_vs_cpu"] = round(rec["mpaths_per_s"] / rec["cpu_baseline"]["value"], 1)
    return rec


def reduction_measure(ctx, w, qv, leaf_bm: int, piece_factor: float, steps: int, max_in_cell: int = 8):
    """Per-cell beam reduction (-gridmax 8, each -gridred type, grid 256) on the headline
    1920x1080 workload: beam-estimator time per step and the reduction grid's build time, next
    to the unreduced structure built the same way."""
    import torch

    from smallupbp_amd import native

    Q = w.queries.count
    out_b = torch.empty((Q, 3), device="cuda")
    rec = {"workload": "bre_beams_1080p beams only, max_beams_in_cell %d, grid 256" % max_in_cell}
    names = ["none", "presample", "offset", "resample_fixed", "resample"]
    for rtype in (-1, 0, 1, 2, 3):
        bm = native.PhotonBeamsEvaluator(ctx)
        torch.cuda.synchronize()
        t = time.perf_counter()
        bm.build(w.beams, leaf_size=leaf_bm, piece_factor=piece_factor,
                 max_beams_in_cell=max_in_cell if rtype >= 0 else 0, reduction_type=max(rtype, 0))
        t_build = time.perf_counter() - t
        el, per = time_eval(ctx, None, bm, qv, None, out_b, steps, 1)
        rec[names[rtype + 1]] = {"build_ms": round(t_build * 1e3, 2),
                                 "kernel_ms": round(float(np.mean([p["kernel_ms_beams"] for p in per])), 3),
                                 "beam_hits": int(per[-1]["beam_hits"])}
        bm.destroy()
    return rec


def calibrate_copy(nbytes: int = 1 << 30):
    """A plain device copy of known size, so PMC runs can scale FETCH_SIZE / WRITE_SIZE."""
    import torch

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("calibration copy: %d bytes read, %d bytes written" % (nbytes, nbytes))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU oracle work (0 = skip)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--order", choices=["sorted", "as_given", "tuned"], default="tuned")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--leaf-size", type=int, default=0)
    ap.add_argument("--piece-factor", type=float, default=0.0)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--photon-waves-per-block", type=int, default=0,
                    help="waves per block of the photon kernel alone (default: --waves-per-block, else tuned)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="ignore smallupbp_amd/tuning_cache.json and use the library defaults")
    ap.add_argument("--batched", type=float, default=1.0,
                    help="also time the 64-instance batched form at this size scale (1 = full size, 0 = skip)")
    ap.add_argument("--no-vm", action="store_true", help="skip the PP3D and SURF vertex-merging measurements")
    ap.add_argument("--no-upbp", action="store_true", help="skip the MIS-weighted (UPBP) measurement")
    ap.add_argument("--no-lighttrace", action="store_true", help="skip the light sub-path tracing measurement")
    ap.add_argument("--no-reduction", action="store_true", help="skip the per-cell beam reduction measurement")
    ap.add_argument("--no-stress", action="store_true", help="skip the 1M x 2M photon-beam stress measurement")
    ap.add_argument("--no-iteration", action="store_true", help="skip the trace + build + eval iteration measurement")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the estimators run on one MI355X and nothing here uses a "
              "second GPU (no collectives, no sharding). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from smallupbp_amd import _abi, native, scenes

    if args.calibrate_copy:
        calibrate_copy()

    t = time.time()
    w = scenes.bre_beams()
    log("workload %s generated in %.1fs" % (w.name, time.time() - t))
    # launch / build parameters: command line, else the autotuner's cached best for this size, else defaults
    from smallupbp_amd import autotune

    tp = autotune.tuned_params(w.photons.count, w.beams.count, w.queries.count)
    tuned = None if args.no_tuning or not tp["cached"] else tp
    leaf_ph = args.leaf_size or (tuned["photon_leaf_size"] if tuned else 0)
    leaf_bm = args.leaf_size or (tuned["beam_leaf_size"] if tuned else 0)
    piece_factor = args.piece_factor or (tuned["piece_factor"] if tuned else 0.0)
    waves = args.waves_per_block or (tuned["beam_waves_per_block"] if tuned else 0)
    waves_ph = args.photon_waves_per_block or args.waves_per_block or (tuned["photon_waves_per_block"] if tuned else 0)
    if args.order == "tuned":
        args.order = "as_given" if tuned and tuned["order"] == _abi.ORDER_AS_GIVEN else "sorted"
    ctx = native.Context(0)
    ctx.set_media(w.media)
    if waves or waves_ph:
        ctx.set_tuning(waves or 4, 0, waves_ph)
    bre = native.EmbreeBre(ctx)
    beams = native.PhotonBeamsEvaluator(ctx)
    torch.cuda.synchronize()
    t = time.perf_counter()
    bre.build(w.photons, leaf_size=leaf_ph)
    t_bp = time.perf_counter() - t
    t = time.perf_counter()
    beams.build(w.beams, leaf_size=leaf_bm, piece_factor=piece_factor)
    t_bb = time.perf_counter() - t
    log("build: photons %.1f ms, beams %.1f ms (%d pieces)" % (t_bp * 1e3, t_bb * 1e3, beams.pieces))
    order = _abi.ORDER_SORTED if args.order == "sorted" else _abi.ORDER_AS_GIVEN
    qv = native.QueryBatchView.to_device(w.queries, order=order)
    Q = w.queries.count
    out_p = torch.empty((Q, 3), device="cuda")
    out_b = torch.empty((Q, 3), device="cuda")

    el, per = time_eval(ctx, bre, beams, qv, out_p, out_b, args.steps, args.warmup)
    ms_step = el / args.steps * 1e3
    value = Q * args.steps / el / 1e6
    ms_ph = float(np.mean([p["kernel_ms_photons"] for p in per]))
    ms_bm = float(np.mean([p["kernel_ms_beams"] for p in per]))
    ms_ot = float(np.mean([p["kernel_ms_other"] for p in per]))
    H_ph = int(per[-1]["photon_hits"])
    H_bm = int(per[-1]["beam_hits"])
    qbytes = alg_bytes_queries(w.queries)
    bytes_ph = qbytes + 28 * H_ph
    bytes_bm = qbytes + 48 * H_bm
    # The photon and beam kernels run concurrently (two streams), so each one's measured duration
    # spans the shared part of the step. The dominant kernel is the one with the larger
    # algorithmic work; its duration is its own launch-to-end time on its stream.
    if bytes_bm >= bytes_ph:
        dom, dom_ms, dom_bytes = "k_eval<beams>", ms_bm, bytes_bm
    else:
        dom, dom_ms, dom_bytes = "k_eval<photons>", ms_ph, bytes_ph
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log("step %.2f ms: photons %.2f ms, beams %.2f ms, sort+clear %.2f ms; hits %d + %d"
        % (ms_step, ms_ph, ms_bm, ms_ot, H_ph, H_bm))
    # each estimator alone (one kernel per call): the per-kernel times without the other's overlap
    n_alone = max(3, args.steps // 2)
    el_p, per_p = time_eval(ctx, bre, None, qv, out_p, None, n_alone, 1)
    el_b, per_b = time_eval(ctx, None, beams, qv, None, out_b, n_alone, 1)
    alone = {"photons_ms_per_step": round(el_p / n_alone * 1e3, 4),
             "photon_kernel_ms": round(float(np.mean([p["kernel_ms_photons"] for p in per_p])), 4),
             "beams_ms_per_step": round(el_b / n_alone * 1e3, 4),
             "beam_kernel_ms": round(float(np.mean([p["kernel_ms_beams"] for p in per_b])), 4),
             "beam_tests_per_hit": round(int(per_b[-1]["beam_candidates"]) / max(1, int(per_b[-1]["beam_hits"])), 3),
             "photon_tests_per_hit": round(int(per_p[-1]["photon_candidates"]) / max(1, int(per_p[-1]["photon_hits"])),
                                           3)}
    log("alone: photons %.2f ms, beams %.2f ms" % (alone["photon_kernel_ms"], alone["beam_kernel_ms"]))

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
            traffic = tj.get(dom)
        except (OSError, ValueError):
            traffic = None

    rec = {
        "metric": "Mqueries/s (BRE + photon beams, 1920x1080 queries x 1M photons + 200k beams)",
        "value": round(value, 3),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox photons, beams and camera rays)",
        "config": {
            "workload": "bre_beams_1080p",
            "queries": Q,
            "photons": w.photons.count,
            "beams": w.beams.count,
            "photon_radius": "lognormal(mu=ln 0.005, sigma=0.4) clamped [5e-4, 0.05]",
            "beam_radius": 0.003,
            "media": "box sigma_s=0.2 sigma_a=0.02 + sphere r=0.3 sigma_s=2.0 sigma_a=0.05",
            "query_order": args.order,
            "beam_pieces": beams.pieces,
            "tuning": {"photon_leaf_size": leaf_ph or "default", "beam_leaf_size": leaf_bm or "default",
                       "beam_piece_factor": piece_factor or "default", "waves_per_block": waves or "default",
                       "photon_waves_per_block": waves_ph or "default",
                       "source": "command line / tuning_cache.json" if tuned else "command line / defaults"},
            "parallelism": "single GPU",
        },
        "kernel_ms": {"photons": round(ms_ph, 4), "beams": round(ms_bm, 4), "sort_and_clear": round(ms_ot, 4),
                      "note": "photon and beam kernels run concurrently on two streams; each time is launch to "
                              "end on its own stream"},
        "candidates": {"photon_tests": int(per[-1]["photon_candidates"]), "beam_tests": int(per[-1]["beam_candidates"]),
                       "nodes": int(per[-1]["nodes_visited"])},
        "alone": alone,
        "hits": {"photon": H_ph, "beam": H_bm,
                 "photon_evals_per_s": round(H_ph / (ms_step * 1e-3), 1),
                 "beam_evals_per_s": round(H_bm / (ms_step * 1e-3), 1)},
        "build_ms": {"photons": round(t_bp * 1e3, 2), "beams": round(t_bb * 1e3, 2)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": traffic,
                     "alg_bytes_per_launch": dom_bytes},
    }

    if not args.no_adjoint:
        ratio, tf, tfa = adjoint_ratio(max(3, args.steps // 2))
        rec["adjoint"] = {"workload": "adjoint_512 (512x512 x 500k photons + 100k beams, L2 loss against the "
                                      "image at sigma_s x 1.1; gradients of 18 medium coefficients, 500k photon "
                                      "and 100k beam powers)",
                          "fwd_ms": round(tf, 3), "fwd_plus_adjoint_ms": round(tfa, 3),
                          "ratio": round(ratio, 3)}

    if args.batched > 0:
        rec["batched"] = batched_ratio(args.batched, max(2, args.steps // 2))

    if not args.no_upbp:
        rec["upbp"] = upbp_measure(max(3, args.steps // 2), min(args.cpu_budget, 5.0), host_cores())

    if not args.no_lighttrace:
        rec["lighttrace"] = light_trace_measure(max(3, args.steps // 2), min(args.cpu_budget, 5.0), host_cores())

    if not args.no_reduction:
        rec["beam_reduction"] = reduction_measure(ctx, w, qv, leaf_bm, piece_factor, max(3, args.steps // 2))

    if not args.no_stress:
        rec["stress"] = stress_measure(max(3, args.steps // 2), min(args.cpu_budget, 10.0), host_cores())

    if not args.no_iteration:
        rec["iteration"] = iteration_measure(max(3, args.steps // 2), ctx, qv, leaf_ph, leaf_bm, piece_factor)
        ctx.set_media(w.media)

    if not args.no_vm:
        from smallupbp_amd import scenes as _sc

        for tech, key in ((_sc.PP3D, "pp3d"), (_sc.SURF, "surf")):
            rec[key] = vm_measure(tech, max(3, args.steps // 2), min(args.cpu_budget, 5.0), host_cores())

    if args.cpu_budget > 0:
        info = host_cpu_info()
        threads = info["threads"]
        mq, n, el_cpu, bs = run_cpu_baseline(w, args.cpu_budget, threads)
        rec["cpu_baseline"] = {
            "value": round(mq, 5), "unit": "Mqueries/s", "cores": threads, "kind": "port",
            "sample": "%d random queries of the same 1920x1080 workload, photons via binned-SAH BVH, beams via "
                      "the reference uniform grid (256 cells, 3D-DDA); %.1f s of eval on %d threads (%s)"
                      % (n, el_cpu, threads, info["model"]),
            "host": {"affinity_cpus": info["affinity"], "os_cpu_count": info["os_cpu_count"],
                     "job_cpu_share": info["job_share"], "model": info["model"],
                     "note": "threads = every CPU of the process's affinity mask within the CPU share the GPU host "
                             "gives this job (OMP_NUM_THREADS)"},
        }
        rec["gpu_vs_cpu"] = round(value / mq, 1)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
