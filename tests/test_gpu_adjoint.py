"""GPU adjoint: gradients of an L2 image loss w.r.t. sigma_s, sigma_a (per medium, RGB) and photon power.

Checked against (a) the analytic CPU gradient of the oracle (FP64 arithmetic on the reference's
FP32 hit sets, which the GPU's equal) and (b) central finite differences (h = 1e-3) of the FP64
oracle forward. Tolerance 1e-3 relative
(to the largest gradient entry of the group; per entry in the SHORT / reduced cases and in
the full-size test). Media are chromatic: with grey media the
minimum-sigma_t channel that normalises SHORT beams switches under a +-h perturbation, and
the estimate has a kink there (the analytic gradient holds that channel fixed).
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from smallupbp_amd import native, scenes

from gpu_helpers import gpu_context, print_case

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _small():
    return scenes.adjoint_scene(40, 40, 6_000, 1_500, seed=21, chromatic=True)


def _oracle_loss(w, media, wp, wb):
    q = scenes.with_media(w, media).queries
    op, _, _ = po.bre_eval(po.OraclePhotons(w.photons), media, q, accel=0, precision=64)
    ob, _, _ = po.beams_eval(po.OracleBeams(w.beams, grid_res=0), media, q, accel=0, precision=64)
    return float((wp * op).sum() + (wb * ob).sum())


def test_adjoint_matches_fp64_gradient_and_fd():
    print_case("adjoint_small")
    w = _small()
    rng = np.random.default_rng(3)
    Q = w.queries.count
    wp = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    wb = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries)
    M = len(w.media)
    gs = torch.zeros((M, 3), device="cuda")
    ga = torch.zeros((M, 3), device="cuda")
    gpp = torch.zeros((w.photons.count, 3), device="cuda")
    gbp = torch.zeros((w.beams.count, 3), device="cuda")
    st = {}
    native.adjoint(ctx, bre, bm, qv, torch.from_numpy(wp).cuda(), torch.from_numpy(wb).cuda(), gs, ga, gpp, gbp,
                   stats=st)
    gs, ga, gpp, gbp = [t.cpu().numpy() for t in (gs, ga, gpp, gbp)]
    ref = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), po.OracleBeams(w.beams, grid_res=0), w.media,
                     w.queries, wp.astype(np.float64), wb.astype(np.float64), fp32_hits=True)
    print("hits", st["photon_hits"], st["beam_hits"])
    print("rel err sigma_s %.2e sigma_a %.2e power %.2e beam power %.2e" % (
        _rel(gs, ref["grad_sigma_s"]), _rel(ga, ref["grad_sigma_a"]), _rel(gpp, ref["grad_photon_power"]),
        _rel(gbp, ref["grad_beam_power"])))
    assert _rel(gs, ref["grad_sigma_s"]) < 1e-3
    assert _rel(ga, ref["grad_sigma_a"]) < 1e-3
    assert _rel(gpp, ref["grad_photon_power"]) < 1e-3
    assert _rel(gbp, ref["grad_beam_power"]) < 1e-3
    # central finite differences of the FP64 forward, h = 1e-3, on the scattering media
    h = 1e-3
    for m in (0, 2):
        for c in range(3):
            for field, g in (("scattering", gs), ("absorption", ga)):
                mp = w.media.copy()
                mm = w.media.copy()
                mp[m][field][c] += h
                mm[m][field][c] -= h
                fd = (_oracle_loss(w, mp, wp, wb) - _oracle_loss(w, mm, wp, wb)) / (2 * h)
                scale = max(np.abs(g).max(), 1e-30)
                assert abs(fd - g[m, c]) / scale < 1e-3, (field, m, c, fd, g[m, c])


def test_autograd_function_l2_loss():
    """torch.autograd.Function wrapper: L2 loss against a target rendered at sigma_s * 1.1."""
    from smallupbp_amd import autograd as ag

    print_case("autograd_l2")
    w = _small()
    target_w = scenes.with_media(w, scenes.scale_scattering(w.media, 1.1))
    tgt_scene = ag.DiffScene(target_w)
    tp, tb = tgt_scene.forward_raw()
    target = (tp + tb).detach()
    scene = ag.DiffScene(w)
    ss, sa = ag.media_params(w.media)
    ss.requires_grad_(True)
    sa.requires_grad_(True)
    power = torch.from_numpy(w.photons.power).cuda().requires_grad_(True)
    op, ob = ag.estimate(scene, ss, sa, power)
    loss = ((op + ob - target) ** 2).sum()
    loss.backward()
    # FD on sigma_s of the box medium, red channel, through the same GPU forward
    h = 1e-3

    def L(delta):
        s2 = ss.detach().clone()
        s2[0, 0] += delta
        a, b = ag.estimate(scene, s2, sa.detach(), power.detach())
        return float(((a + b - target) ** 2).double().sum())

    fd = (L(h) - L(-h)) / (2 * h)
    g = float(ss.grad[0, 0])
    print("dL/dsigma_s[box,r]: adjoint %.6e fd %.6e" % (g, fd))
    assert abs(fd - g) <= 2e-2 * abs(g) + 1e-12
    # photon power: linear, FD exact up to FP32 rounding
    assert torch.isfinite(power.grad).all()


def test_adjoint_gradient_of_negative_coefficients_carries_their_sign():
    """The media use |sigma|: the GPU gradient with respect to a negative input coefficient is
    the oracle's (whose sign is pinned on the CPU by central differences), sign included."""
    print_case("adjoint_negative_sigma")
    w = _small()
    media = w.media.copy()
    media[0]["absorption"][1] = -media[0]["absorption"][1]
    media[2]["scattering"][0] = -media[2]["scattering"][0]
    w = scenes.with_media(w, media)
    rng = np.random.default_rng(5)
    Q = w.queries.count
    wp = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    wb = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    ctx = gpu_context(media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries)
    M = len(media)
    for dev in (True, False):
        gs = torch.zeros((M, 3), device="cuda") if dev else np.zeros((M, 3), np.float32)
        ga = torch.zeros((M, 3), device="cuda") if dev else np.zeros((M, 3), np.float32)
        gp = torch.from_numpy(wp).cuda() if dev else wp
        gb = torch.from_numpy(wb).cuda() if dev else wb
        q = qv if dev else native.QueryBatchView.host(w.queries)
        native.adjoint(ctx, bre, bm, q, gp, gb, gs, ga, None, None)
        if dev:
            gs, ga = gs.cpu().numpy(), ga.cpu().numpy()
        ref = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), po.OracleBeams(w.beams, grid_res=0), media,
                         w.queries, wp.astype(np.float64), wb.astype(np.float64), fp32_hits=True)
        for got, want, m, c in ((ga, ref["grad_sigma_a"], 0, 1), (gs, ref["grad_sigma_s"], 2, 0)):
            assert np.sign(got[m, c]) == np.sign(want[m, c]) != 0
            assert abs(got[m, c] - want[m, c]) <= 1e-3 * abs(want[m, c]), (dev, m, c, got[m, c], want[m, c])
        assert _rel(gs, ref["grad_sigma_s"]) < 1e-3
        assert _rel(ga, ref["grad_sigma_a"]) < 1e-3


@pytest.mark.parametrize("short,reduced", [(True, False), (False, True), (True, True)])
def test_adjoint_short_queries_and_reduced_beams_match_fp64_gradient(short, reduced):
    """SHORT query beams (in-segment transmittance over its minimum channel, earlier segments over
    their pass-through pdfs) and a beam grid reduced to 2 beams per crowded cell: the GPU adjoint
    equals the FP64 oracle gradient (itself pinned by central differences on the CPU), each of the
    12 scattering-medium entries to 1e-3 of itself plus 1e-5 of the largest (FP32 accumulation of
    near-cancelling entries), and the photon and beam power gradients to 1e-3."""
    import dataclasses

    print_case("adjoint_short_reduced", short=short, reduced=reduced)
    w = _small()
    if short:
        q = w.queries
        w = scenes.with_media(dataclasses.replace(w, queries=scenes.Queries(q.origin, q.direction, q.seg_offsets,
                                                                          q.segments, scenes.SHORT_BEAM)), w.media)
    rng = np.random.default_rng(6)
    Q = w.queries.count
    wp = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    wb = rng.uniform(-1, 1, (Q, 3)).astype(np.float32)
    ctx = gpu_context(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    if reduced:
        bm.build(w.beams, max_beams_in_cell=2, reduction_type=scenes.REDUCE_RESAMPLE_FIXED, grid_resolution=8,
                 reduction_seed=3)
        obm = po.OracleBeams(w.beams, grid_res=8).set_reduction(2, scenes.REDUCE_RESAMPLE_FIXED, 3)
    else:
        bm.build(w.beams)
        obm = po.OracleBeams(w.beams, grid_res=0)
    qv = native.QueryBatchView.to_device(w.queries)
    M = len(w.media)
    gs = torch.zeros((M, 3), device="cuda")
    ga = torch.zeros((M, 3), device="cuda")
    gpp = torch.zeros((w.photons.count, 3), device="cuda")
    gbp = torch.zeros((w.beams.count, 3), device="cuda")
    native.adjoint(ctx, bre, bm, qv, torch.from_numpy(wp).cuda(), torch.from_numpy(wb).cuda(), gs, ga, gpp, gbp)
    gs, ga, gpp, gbp = [t.cpu().numpy().astype(np.float64) for t in (gs, ga, gpp, gbp)]
    ref = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), obm, w.media, w.queries, wp.astype(np.float64),
                     wb.astype(np.float64), fp32_hits=True)
    big = max(np.abs(ref["grad_sigma_s"]).max(), np.abs(ref["grad_sigma_a"]).max())
    for got, want, name in ((gs, ref["grad_sigma_s"], "sigma_s"), (ga, ref["grad_sigma_a"], "sigma_a")):
        for m in (0, 2):
            for c in range(3):
                assert abs(got[m, c] - want[m, c]) <= 1e-3 * abs(want[m, c]) + 1e-5 * big, (name, m, c, got[m, c],
                                                                                              want[m, c])
    assert _rel(gpp, ref["grad_photon_power"]) < 1e-3
    assert _rel(gbp, ref["grad_beam_power"]) < 1e-3
    if reduced:
        plain = po.adjoint(po.OraclePhotons(w.photons, build_bvh=False), po.OracleBeams(w.beams, grid_res=0), w.media,
                           w.queries, wp.astype(np.float64), wb.astype(np.float64), fp32_hits=True)
        assert np.abs(plain["grad_beam_power"] - ref["grad_beam_power"]).max() > 0  # the reduction acted
