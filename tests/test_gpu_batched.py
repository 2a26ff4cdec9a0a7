"""Batched persistent evaluation of many uneven instances (device work queue, hipGraph replay).

Each instance's result must equal its own plain evaluation (same kernels, same
query bundles), on eager runs and on every graph replay.
"""
import numpy as np
import pytest
import torch

from smallupbp_amd import _abi, native, scenes

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


def _two_instance_plan(ws, media, bases):
    ctx = native.Context(0)
    ctx.set_media(media)
    keep, inst = [], []
    for w, base in zip(ws, bases):
        bre = native.EmbreeBre(ctx)
        bre.build(w.photons)
        bm = native.PhotonBeamsEvaluator(ctx)
        bm.build(w.beams)
        qv = native.QueryBatchView.to_device(w.queries)
        op = torch.full((w.queries.count, 3), -1.0, device="cuda")
        ob = torch.full((w.queries.count, 3), -1.0, device="cuda")
        keep.append((bre, bm, qv, op, ob))
        inst.append((bre, bm, qv, base, op, ob))
    return ctx, keep, native.BatchPlan(ctx, inst)


def test_batched_rejects_a_medium_id_outside_the_instances_range():
    """An instance's media are the table entries from its media_base on: a segment medium id
    that is valid for the whole table but not for the instance (media_base + id past the end)
    fails the run with UPBP_EINVAL instead of reading past the media table."""
    print_case("batched_medium_range")
    ws = _instances(2, 0.005, seed=11)
    media = np.concatenate([w.media for w in ws])  # 2 media per instance, 4 in all
    w1 = ws[1]
    seg = w1.queries.segments.copy()
    seg["medium_id"][0] = 2  # < 4 (the table) but >= 4 - 2 (the instance's range)
    bad = scenes.Workload(w1.name, w1.media, w1.photons, w1.beams,
                          scenes.Queries(w1.queries.origin, w1.queries.direction, w1.queries.seg_offsets, seg,
                                         w1.queries.beam_type))
    ctx, keep, plan = _two_instance_plan([ws[0], bad], media, [0, 2])
    for use_graph in (False, True):
        with pytest.raises(native.UpbpError) as e:
            plan.run(use_graph=use_graph)
        assert e.value.status == _abi.UPBP_EINVAL
    # the same plan with valid ids runs
    ctx2, keep2, plan2 = _two_instance_plan(ws, media, [0, 2])
    plan2.run(use_graph=True)
    torch.cuda.synchronize()


def test_batched_graph_follows_a_new_media_table():
    """upbp_set_media between graph runs: the next run re-validates the instances and captures
    again, so it reads the new table (grown here, which moves the device buffer) and equals an
    eager run; a table too small for an instance's media_base is rejected."""
    print_case("batched_media_regrowth")
    ws = _instances(2, 0.005, seed=12)
    media = np.concatenate([w.media for w in ws])
    ctx, keep, plan = _two_instance_plan(ws, media, [0, 2])
    plan.run(use_graph=True)
    torch.cuda.synchronize()
    first = [(op.cpu().numpy(), ob.cpu().numpy()) for _, _, _, op, ob in keep]
    # grow the table: instance 1 now reads new coefficients at the same ids
    grown = np.concatenate([media, media[:2]] * 8)
    grown[2]["scattering"] = grown[2]["scattering"] * np.float32(1.5)
    ctx.set_media(grown)
    plan.run(use_graph=True)
    torch.cuda.synchronize()
    graph_out = [(op.cpu().numpy(), ob.cpu().numpy()) for _, _, _, op, ob in keep]
    plan.run(use_graph=False)
    torch.cuda.synchronize()
    eager_out = [(op.cpu().numpy(), ob.cpu().numpy()) for _, _, _, op, ob in keep]
    for (gp, gb), (ep, eb) in zip(graph_out, eager_out):
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    assert np.array_equal(graph_out[0][0], first[0][0])  # instance 0's media did not change
    assert not np.array_equal(graph_out[1][0], first[1][0])  # instance 1's sigma_s did
    ctx.set_media(media[:2])
    with pytest.raises(native.UpbpError) as e:
        plan.run(use_graph=True)
    assert e.value.status == _abi.UPBP_EINVAL
