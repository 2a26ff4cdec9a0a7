"""GPU parity at the full sizes of the benchmark configurations.

The GPU evaluates every query of the workload with the same build and launch parameters
the benchmark uses; a seeded random sample of those queries is then evaluated by the CPU
oracle (photons through its BVH or by brute force, beams by brute force) and compared:
hit signatures identical, radiance within REL_TOL = 1e-4 of the FP32 oracle. The one
documented exception is a beam end tie (gpu_helpers.check_beam_ties): the reference accepts a
nearly perpendicular pair whose closest point lies beyond the beam's end through the rounding
of its t2 = n / q, and the GPU, which visits only pieces near the beam, misses it. Properties
that hold over all queries, where the oracle cannot follow at full size, are checked too:
the kernels' hit counters equal the sum of the per-query signature counts, and the
batched plan finds exactly as many hits as the per-instance evaluations.

Workloads (smallupbp_amd/scenes.py):
  bre_beams_1080p   1920x1080 queries x 1M photons + 200k SHORT beams, box + sphere media
  beams_stress      1M query beams x 2M beams (50 % SHORT, 50 % LONG), one medium
  batched_64        64 uneven instances at full size (10k-2M photons, 16k-2M queries each)
  adjoint_512       512x512 queries x 500k photons + 100k beams, L2 loss against sigma_s x 1.1
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from smallupbp_amd import autotune, native, scenes

from gpu_helpers import REL_TOL, beam_mismatch_report, check_beam_ties, gpu_context, print_case, rel_err

pytestmark = pytest.mark.gpu


def _sample(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, min(k, n), replace=False))


def _tuned_context(w):
    """Context, structures and device queries with the benchmark's tuned parameters."""
    p = autotune.tuned_params(w.photons.count if w.photons is not None else 0,
                              w.beams.count if w.beams is not None else 0, w.queries.count)
    ctx = gpu_context(w.media)
    if p["beam_waves_per_block"] or p["photon_waves_per_block"]:
        ctx.set_tuning(p["beam_waves_per_block"] or 4, 0, p["photon_waves_per_block"])
    bre = bm = None
    if w.photons is not None:
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons, leaf_size=p["photon_leaf_size"])
    if w.beams is not None:
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams, leaf_size=p["beam_leaf_size"], piece_factor=p["piece_factor"])
    qv = native.QueryBatchView.to_device(w.queries, order=p["order"])
    return ctx, bre, bm, qv, p


def _check_sample(name, out, sig, ref, rsig, cnt, explain=None, ties=None):
    """Identical hit signatures (beams: except documented end ties, ties(sig, rsig) -> count),
    radiance within REL_TOL of the FP32 oracle on the queries whose hit sets agree."""
    same = np.all(sig == rsig, axis=1)
    if not same.all() and explain is not None:
        print(explain(sig, rsig), flush=True)
    if ties is None:
        assert same.all(), "%s hit sets differ from the oracle on the sample" % name
    else:
        n = ties(sig, rsig)
        print("%s: %d queries with end ties (%d pairs)" % (name, int((~same).sum()), n), flush=True)
        assert (~same).mean() < 0.01
    assert int(sig[same, 0].astype(np.int64).sum()) == int(rsig[same, 0].astype(np.int64).sum())
    e = rel_err(out[same], ref[same])
    print("%s: %d sampled queries, %d hits, max rel err %.3g" % (name, len(sig), int(cnt[0]), e.max()), flush=True)
    assert e.max() <= REL_TOL


@pytest.mark.timeout(600)
def test_bre_beams_1080p_full_size():
    """1920x1080 queries x 1M photons + 200k beams: the headline workload with its tuned
    leaf sizes, piece length and launch shape; 20,000 sampled queries against the oracle's
    photon BVH (FP32 and FP64), 2,000 against its beam brute force."""
    print_case("bre_beams_1080p_full")
    w = scenes.bre_beams()
    ctx, bre, bm, qv, p = _tuned_context(w)
    st = {}
    op, ob = native.eval_batch(ctx, bre, bm, qv, stats=st)
    op, ob = op.cpu().numpy(), ob.cpu().numpy()
    sp, sb = native.hit_signatures(ctx, bre, bm, qv)
    print("hits: %d photon, %d beam; %d beam pieces; params %s" % (st["photon_hits"], st["beam_hits"], bm.pieces, p),
          flush=True)
    # over all 2,073,600 queries: the forward kernels' counters equal the signature kernels' counts
    assert st["photon_hits"] == int(sp[:, 0].astype(np.int64).sum())
    assert st["beam_hits"] == int(sb[:, 0].astype(np.int64).sum())
    assert np.isfinite(op).all() and np.isfinite(ob).all()
    idx = _sample(w.queries.count, 20_000, 101)
    qs = w.queries.subset(idx)
    oph = po.OraclePhotons(w.photons)
    ref, rsig, cnt = po.bre_eval(oph, w.media, qs, accel=1, signatures=True)
    _check_sample("photons", op[idx], sp[idx], ref, rsig, cnt)
    ref64, _, _ = po.bre_eval(oph, w.media, qs, accel=1, precision=64)
    e64 = rel_err(op[idx], ref64)
    print("photons vs FP64: median %.3g p99.9 %.3g" % (np.median(e64), np.percentile(e64, 99.9)), flush=True)
    assert np.median(e64) < 1e-5
    jdx = _sample(w.queries.count, 2_000, 102)
    qs = w.queries.subset(jdx)
    ref, rsig, cnt = po.beams_eval(po.OracleBeams(w.beams, grid_res=0), w.media, qs, accel=0, signatures=True)
    _check_sample("beams", ob[jdx], sb[jdx], ref, rsig, cnt,
                  lambda g, r: beam_mismatch_report(w, qs, g, r, piece_factor=p["piece_factor"] or 8.0),
                  lambda g, r: check_beam_ties(qs, w.beams, g, r))


@pytest.mark.timeout(600)
def test_subtree_leaves_add_only_reference_hits():
    """The 1080p beams with the built leaves alone (frustum_leaf 1) and with frustum waves taking
    subtrees of up to 64 items as one leaf (the default): the larger subtree boxes reach end-tie
    pairs that the leaf boxes miss, so the hit sets may differ, but only by pairs the reference
    accepts. Every differing query is compared with the oracle's brute force: the default's hit
    set equals the oracle's up to end ties, and it holds every hit of the leaf-only run."""
    print_case("subtree_leaves_1080p")
    w = scenes.bre_beams()
    w = scenes.Workload(w.name, w.media, None, w.beams, w.queries, w.sphere, w.box_medium, w.vacuum_medium, w.seed)
    ctx, _, bm, qv, p = _tuned_context(w)
    wpb = p["beam_waves_per_block"] or 4
    sigs, hits = {}, {}
    for fl in (1, 64):
        ctx.set_tuning(wpb, 0, 0, fl)
        st = {}
        native.eval_batch(ctx, None, bm, qv, stats=st)
        hits[fl] = st["beam_hits"]
        sigs[fl] = native.hit_signatures(ctx, None, bm, qv)[1]
    diff = np.nonzero(np.any(sigs[1] != sigs[64], axis=1))[0]
    print("beam hits: leaves only %d, subtrees of 64 %d; %d queries differ" % (hits[1], hits[64], len(diff)),
          flush=True)
    assert hits[64] >= hits[1]
    assert np.all(sigs[64][:, 0] >= sigs[1][:, 0])
    assert len(diff) < 1000
    if len(diff):
        qs = w.queries.subset(diff)
        _, rsig, _ = po.beams_eval(po.OracleBeams(w.beams, grid_res=0), w.media, qs, accel=0, signatures=True)
        for fl in (1, 64):
            n = check_beam_ties(qs, w.beams, sigs[fl][diff], rsig)
            print("frustum_leaf %d: %d end-tie pairs missed on the differing queries" % (fl, n), flush=True)
        # the default misses fewer of the reference's pairs there than the leaf-only run
        assert (sigs[64][diff, 0].astype(np.int64) - rsig[:, 0].astype(np.int64)).sum() > \
            (sigs[1][diff, 0].astype(np.int64) - rsig[:, 0].astype(np.int64)).sum()


@pytest.mark.timeout(600)
def test_beams_stress_full_size():
    """1M query beams x 2M photon beams (half SHORT, half LONG): 2,000 sampled queries
    against the oracle's brute force; their total hit count equals the oracle's."""
    print_case("beams_stress_full")
    w = scenes.beams_stress()
    ctx, _, bm, qv, p = _tuned_context(w)
    st = {}
    _, ob = native.eval_batch(ctx, None, bm, qv, stats=st)
    ob = ob.cpu().numpy()
    _, sb = native.hit_signatures(ctx, None, bm, qv)
    print("beam hits %d (%.0f per query), %d pieces" % (st["beam_hits"], st["beam_hits"] / w.queries.count, bm.pieces),
          flush=True)
    assert st["beam_hits"] == int(sb[:, 0].astype(np.int64).sum())
    assert np.isfinite(ob).all()
    idx = _sample(w.queries.count, 2_000, 103)
    qs = w.queries.subset(idx)
    ref, rsig, cnt = po.beams_eval(po.OracleBeams(w.beams, grid_res=0), w.media, qs, accel=0, signatures=True)
    _check_sample("beams", ob[idx], sb[idx], ref, rsig, cnt,
                  lambda g, r: beam_mismatch_report(w, qs, g, r, piece_factor=p["piece_factor"] or 8.0),
                  lambda g, r: check_beam_ties(qs, w.beams, g, r))


@pytest.mark.timeout(900)
def test_batched_64_full_size():
    """64 uneven instances at their stated sizes in one persistent plan, replayed from a
    hipGraph: per instance 300 sampled queries against the oracle (values; the hit
    signatures of the same queries through the per-instance path equal the oracle's), and
    the plan's total hit counts equal the per-instance evaluations' over all queries."""
    print_case("batched_64_full")
    ws = scenes.batched_instances(64, seed=4, scale=1.0)
    Q = sum(w.queries.count for w in ws)
    print("%d instances, %d photons, %d beams, %d queries" % (
        len(ws), sum(w.photons.count for w in ws), sum(w.beams.count for w in ws), Q), flush=True)
    media = np.concatenate([w.media for w in ws])
    ctx = native.Context(0)
    ctx.set_media(media)
    inst, per = [], []
    for i, w in enumerate(ws):
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons)
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams)
        qv = native.QueryBatchView.to_device(w.queries)
        op = torch.full((w.queries.count, 3), -1.0, device="cuda")
        ob = torch.full((w.queries.count, 3), -1.0, device="cuda")
        inst.append((bre, bm, qv, 2 * i, op, ob))
    plan = native.BatchPlan(ctx, inst)
    st = {}
    plan.run(use_graph=True)
    plan.run(use_graph=True, stats=st)
    torch.cuda.synchronize()
    # per-instance evaluations over all queries (own context: media base 0)
    hp = hb = 0
    for w, (bre, bm, qv, _, _, _) in zip(ws, inst):
        pctx = gpu_context(w.media)
        pb = native.EmbreeBre(pctx)
        pb.build(w.photons)
        pm = native.PhotonBeamsEvaluator(pctx)
        pm.build(w.beams)
        s1 = {}
        native.eval_batch(pctx, pb, pm, native.QueryBatchView.to_device(w.queries), stats=s1)
        hp += s1["photon_hits"]
        hb += s1["beam_hits"]
        per.append((pctx, pb, pm))
    print("plan hits %d + %d, per-instance %d + %d" % (st["photon_hits"], st["beam_hits"], hp, hb), flush=True)
    assert st["photon_hits"] == hp and st["beam_hits"] == hb
    worst_p = worst_b = 0.0
    tied = 0
    for i, (w, (_, _, _, _, op, ob), (pctx, pb, pm)) in enumerate(zip(ws, inst, per)):
        idx = _sample(w.queries.count, 300, 200 + i)
        qs = w.queries.subset(idx)
        t = torch.from_numpy(idx).cuda()
        gp = op.index_select(0, t).cpu().numpy()
        gb = ob.index_select(0, t).cpu().numpy()
        sp, sb = native.hit_signatures(pctx, pb, pm, native.QueryBatchView.host(qs))
        rp, rsp, _ = po.bre_eval(po.OraclePhotons(w.photons, build_bvh=False), w.media, qs, accel=0, signatures=True)
        rb, rsb, _ = po.beams_eval(po.OracleBeams(w.beams, grid_res=0), w.media, qs, accel=0, signatures=True)
        assert np.array_equal(sp, rsp), "instance %d: photon hit sets differ" % i
        same = np.all(sb == rsb, axis=1)
        if not same.all():
            tied += check_beam_ties(qs, w.beams, sb, rsb)
        ep, eb = rel_err(gp, rp).max(), rel_err(gb[same], rb[same]).max()
        worst_p, worst_b = max(worst_p, ep), max(worst_b, eb)
        assert ep <= REL_TOL and eb <= REL_TOL, (i, ep, eb)
    print("batched vs oracle on 64 x 300 queries: max rel err %.3g photons, %.3g beams; %d beam end ties"
          % (worst_p, worst_b, tied), flush=True)


def _l2_setup():
    """adjoint_512 with chromatic media (a unique minimum sigma_t channel, so central
    differences do not straddle the switch of the SHORT-beam normalisation channel that grey
    media sit on) and its target image at sigma_s x 1.1."""
    w = scenes.adjoint_scene(chromatic=True)
    tgt = scenes.with_media(w, scenes.scale_scattering(w.media, 1.1))
    return w, tgt


def _forward(w, bre, bm, media):
    ctx = bre.ctx
    ctx.set_media(media)
    qv = native.QueryBatchView.to_device(scenes.with_media(w, media).queries)
    op, ob = native.eval_batch(ctx, bre, bm, qv)
    return (op + ob), qv


@pytest.mark.timeout(600)
def test_adjoint_512_full_size():
    """512x512 x (500k photons + 100k beams), loss = sum_q |L_q - T_q|^2 with T rendered at
    sigma_s x 1.1. (a) The GPU adjoint on 2,000 sampled queries equals the oracle gradient of
    the same queries (FP64 arithmetic on the reference's FP32 hit sets): each of the 12 medium
    entries to 1e-3 of itself, every photon-power entry to 1e-3 of itself plus 1e-5 of the sum of
    its terms' magnitudes (FP32 accumulation of terms of both signs). (b) Central differences
    (h = 1e-3) of the GPU forward over all 262,144 queries equal the GPU adjoint's gradient, each
    of the 12 entries to 1e-3."""
    print_case("adjoint_512_full")
    w, tgt = _l2_setup()
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    T, _ = _forward(tgt, bre, bm, tgt.media)
    L, qv = _forward(w, bre, bm, w.media)
    g = (2.0 * (L - T)).contiguous()
    M = len(w.media)
    gs = torch.zeros((M, 3), device="cuda")
    ga = torch.zeros((M, 3), device="cuda")
    gpp = torch.zeros((w.photons.count, 3), device="cuda")
    native.adjoint(ctx, bre, bm, qv, g, g, gs, ga, gpp, None)
    gs, ga = gs.cpu().numpy().astype(np.float64), ga.cpu().numpy().astype(np.float64)
    media_idx = [0, 2]  # box and sphere; medium 1 is the clear medium outside (|sigma| = 0 is a kink)

    # (b) central differences of the GPU forward over every query
    h = 1e-3

    def loss(media):
        Lm, _ = _forward(w, bre, bm, media)
        return float(((Lm.double() - T.double()) ** 2).sum())

    worst = 0.0
    for m in media_idx:
        for field, G in (("scattering", gs), ("absorption", ga)):
            for c in range(3):
                mp, mm = w.media.copy(), w.media.copy()
                mp[m][field][c] += h
                mm[m][field][c] -= h
                fd = (loss(mp) - loss(mm)) / (2 * h)
                err = abs(fd - G[m, c]) / abs(G[m, c])
                worst = max(worst, err)
                print("d loss / d %s[%d][%d]: adjoint %.7e fd %.7e rel %.2e" % (field, m, c, G[m, c], fd, err),
                      flush=True)
                assert err <= 1e-3, (field, m, c, fd, G[m, c])
    print("worst FD relative error over the 12 entries: %.2e" % worst, flush=True)

    # (a) the oracle gradient (FP64 arithmetic on the reference's FP32 hit sets) on a query sample,
    # with the GPU adjoint on the same sample
    ctx.set_media(w.media)
    idx = _sample(w.queries.count, 2_000, 104)
    qs = w.queries.subset(idx)
    gsub = g.index_select(0, torch.from_numpy(idx).cuda()).contiguous()
    gs2 = torch.zeros((M, 3), device="cuda")
    ga2 = torch.zeros((M, 3), device="cuda")
    gpp2 = torch.zeros((w.photons.count, 3), device="cuda")
    native.adjoint(ctx, bre, bm, native.QueryBatchView.to_device(qs), gsub, gsub, gs2, ga2, gpp2, None)
    gh = gsub.cpu().numpy().astype(np.float64)
    ref = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), po.OracleBeams(w.beams, grid_res=0), w.media, qs,
                     gh, gh, fp32_hits=True)
    gs2, ga2, gpp2 = (t.cpu().numpy().astype(np.float64) for t in (gs2, ga2, gpp2))
    for m in media_idx:
        for c in range(3):
            for got, want, name in ((gs2, ref["grad_sigma_s"], "sigma_s"), (ga2, ref["grad_sigma_a"], "sigma_a")):
                assert abs(got[m, c] - want[m, c]) <= 1e-3 * abs(want[m, c]), (name, m, c, got[m, c], want[m, c])
    rp = ref["grad_photon_power"]
    # a photon's power gradient sums w_q * b_q over the sampled queries that hit it, with upstream
    # gradients w_q = 2 (L_q - T_q) of either sign: its FP32 accumulation error scales with the sum
    # of |terms|, which the same oracle gradient with |w_q| gives
    mag = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), None, w.media, qs, np.abs(gh), None,
                     fp32_hits=True)["grad_photon_power"]
    hit = mag.max(1) > 0
    print("photon-power gradient: %d photons hit by the sample" % int(hit.sum()), flush=True)
    assert hit.sum() > 1000
    assert np.all(np.abs(gpp2[~hit]) == 0)
    excess = np.abs(gpp2[hit] - rp[hit]) - (1e-3 * np.abs(rp[hit]) + 1e-5 * mag[hit])
    rel = np.abs(gpp2[hit] - rp[hit]) / np.maximum(np.abs(rp[hit]), 1e-30)
    print("photon-power gradient: median rel err %.2e, entries above 1e-3 relative %d (all within 1e-5 of their "
          "|term| sums: %s)" % (np.median(rel), int((rel > 1e-3).sum()), bool((excess <= 0).all())), flush=True)
    assert (excess <= 0).all()
