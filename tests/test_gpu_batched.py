"""Batched persistent evaluation of many uneven instances (device work queue, hipGraph replay).

Each instance's result must equal its own plain evaluation (same kernels, same
query bundles), on eager runs and on every graph replay.
"""
import numpy as np
import pytest
import torch

from smallupbp_amd import native, scenes

from gpu_helpers import print_case

pytestmark = pytest.mark.gpu


def _instances(n, scale, seed=4):
    return scenes.batched_instances(n_instances=n, seed=seed, scale=scale, max_queries=2_000_000)


def _plain_reference(w):
    ctx = native.Context(0)
    ctx.set_media(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    bm = native.PhotonBeamsEvaluator(ctx)
    bm.build(w.beams)
    qv = native.QueryBatchView.to_device(w.queries)
    op, ob = native.eval_batch(ctx, bre, bm, qv)
    return op.cpu().numpy(), ob.cpu().numpy()


def test_batched_plan_equals_per_instance_eval():
    print_case("batched", n=12, scale=0.01)
    ws = _instances(12, 0.01)
    media = np.concatenate([w.media for w in ws])
    ctx = native.Context(0)
    ctx.set_media(media)
    keep, inst = [], []
    for i, w in enumerate(ws):
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons)
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams)
        qv = native.QueryBatchView.to_device(w.queries)
        op = torch.full((w.queries.count, 3), -1.0, device="cuda")
        ob = torch.full((w.queries.count, 3), -1.0, device="cuda")
        keep.append((bre, bm, qv, op, ob))
        inst.append((bre, bm, qv, 2 * i, op, ob))
    plan = native.BatchPlan(ctx, inst)
    refs = [_plain_reference(w) for w in ws]
    for use_graph in (False, True, True, True):
        for _, _, _, op, ob in keep:
            op.fill_(-1.0)
            ob.fill_(-1.0)
        torch.cuda.synchronize()
        st = {}
        plan.run(use_graph=use_graph, stats=st)
        # eager work between replays
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        assert st["photon_hits"] > 0 and st["beam_hits"] > 0
        for (rp, rb), (_, _, _, op, ob) in zip(refs, keep):
            assert np.array_equal(op.cpu().numpy(), rp)
            assert np.array_equal(ob.cpu().numpy(), rb)
        del junk


def test_batched_plan_rejects_host_queries():
    print_case("batched_host_queries")
    w = _instances(1, 0.005)[0]
    ctx = native.Context(0)
    ctx.set_media(w.media)
    bre = native.EmbreeBre(ctx)
    bre.build(w.photons)
    qv = native.QueryBatchView.host(w.queries)
    op = torch.zeros((w.queries.count, 3), device="cuda")
    with pytest.raises(native.UpbpError):
        native.BatchPlan(ctx, [(bre, None, qv, 0, op, None)])


def test_batched_plan_owns_structure_copies_and_mixed_instances():
    # the plan copies the structures: they may be destroyed after the plan is made;
    # instances may carry photons only, beams only, or both
    print_case("batched_copies", n=5, scale=0.01)
    ws = _instances(5, 0.01, seed=9)
    media = np.concatenate([w.media for w in ws])
    ctx = native.Context(0)
    ctx.set_media(media)
    built, inst, outs = [], [], []
    for i, w in enumerate(ws):
        bre = None if i == 1 else native.EmbreeBre(ctx)
        bm = None if i == 0 else native.PhotonBeamsEvaluator(ctx)
        if bre is not None:
            bre.build(w.photons)
        if bm is not None:
            bm.build(w.beams)
        qv = native.QueryBatchView.to_device(w.queries)
        op = torch.full((w.queries.count, 3), -1.0, device="cuda") if bre is not None else None
        ob = torch.full((w.queries.count, 3), -1.0, device="cuda") if bm is not None else None
        built.append((bre, bm))
        inst.append((bre, bm, qv, 2 * i, op, ob))
        outs.append((op, ob))
    plan = native.BatchPlan(ctx, inst)
    refs = [_plain_reference(w) for w in ws]
    for bre, bm in built:
        if bre is not None:
            bre.destroy()
        if bm is not None:
            bm.destroy()
    for use_graph in (False, True, True):
        plan.run(use_graph=use_graph)
        torch.cuda.synchronize()
        for (rp, rb), (op, ob) in zip(refs, outs):
            if op is not None:
                assert np.array_equal(op.cpu().numpy(), rp)
            if ob is not None:
                assert np.array_equal(ob.cpu().numpy(), rb)
