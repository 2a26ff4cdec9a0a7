"""torch.autograd.Function over the forward (upbp_eval) and adjoint (upbp_adjoint) kernels.

Differentiable parameters: per-medium scattering and absorption (RGB) and
per-photon power. Photon/beam geometry and query rays are fixed. Segment
attenuations are recomputed from the current coefficients on every forward,
exp(-(sigma_a + sigma_s) * (dist_max - dist_min)), which is what the reference's
Scene::Intersect stores for a segment (SmallUPBP/src/Scene/Scene.hxx:542), and so are
the pass-through ray-sampling pdfs exp(-sigma_t_min * (dist_max - dist_min)) that SHORT
query beams divide by (RaySamplePdf, SmallUPBP/src/Scene/Media.hxx:294-323).
"""
from __future__ import annotations

import numpy as np
import torch

from . import native, scenes


class DiffScene:
    """Built structures plus device-resident queries of one differentiable workload."""

    def __init__(self, w: scenes.Workload, device: int = 0):
        self.w = w
        self.ctx = native.Context(device)
        self.ctx.set_media(w.media)
        self.bre = None
        self.beams = None
        if w.photons is not None:
            self.bre = native.EmbreeBre(self.ctx)
            self.bre.build(w.photons)
        if w.beams is not None:
            self.beams = native.PhotonBeamsEvaluator(self.ctx)
            self.beams.build(w.beams)
        self.qv = native.QueryBatchView.to_device(w.queries)
        seg = self.qv._keep[3]  # [S,8] float32 view of the segment records
        self.seg = seg
        self.seg_len = (seg[:, 1] - seg[:, 0]).clone()
        self.seg_mid = seg[:, 7].contiguous().view(torch.int32).long()
        self.base_media = w.media.copy()

    def set_params(self, sigma_s: torch.Tensor, sigma_a: torch.Tensor, photon_power: torch.Tensor | None):
        m = self.base_media.copy()
        m["scattering"] = sigma_s.detach().float().cpu().numpy()
        m["absorption"] = sigma_a.detach().float().cpu().numpy()
        self.ctx.set_media(m)
        ss = sigma_s.detach().abs().float().to(self.seg.device)
        st = ss + sigma_a.detach().abs().float().to(self.seg.device)
        att = torch.exp(-st[self.seg_mid] * self.seg_len[:, None])
        self.seg[:, 4:7] = att
        # pass-through pdf exp(-sigma_t_min len) of media that sample distances (min over positive channels)
        mn = torch.where(st > 0, st, torch.full_like(st, float("inf"))).min(dim=1).values
        samplable = torch.isfinite(mn) & (ss > 0).any(dim=1)
        mn = torch.where(samplable, mn, torch.zeros_like(mn))
        pdf = torch.clamp(torch.exp(-mn[self.seg_mid] * self.seg_len), min=1e-35)
        pdf = torch.where(samplable[self.seg_mid], pdf, torch.ones_like(pdf))
        self.seg[:, 2] = pdf
        self.seg[:, 3] = pdf
        if photon_power is not None and self.bre is not None:
            self.bre.set_power(photon_power.detach().float().contiguous())
        torch.cuda.synchronize()

    def forward_raw(self):
        return native.eval_batch(self.ctx, self.bre, self.beams, self.qv)


class VolumetricEstimate(torch.autograd.Function):
    """(sigma_s [M,3], sigma_a [M,3], photon_power [N,3]) -> (photon radiance [Q,3], beam radiance [Q,3])."""

    @staticmethod
    def forward(fctx, scene: DiffScene, sigma_s, sigma_a, photon_power):
        scene.set_params(sigma_s, sigma_a, photon_power)
        out_p, out_b = scene.forward_raw()
        fctx.scene = scene
        fctx.save_for_backward(sigma_s, sigma_a, photon_power)
        q = scene.qv.count
        zeros = torch.zeros((q, 3), device="cuda")
        return (out_p if out_p is not None else zeros), (out_b if out_b is not None else zeros)

    @staticmethod
    def backward(fctx, g_p, g_b):
        scene = fctx.scene
        sigma_s, sigma_a, photon_power = fctx.saved_tensors
        # the structures are shared: restore this call's parameters before re-traversing
        scene.set_params(sigma_s, sigma_a, photon_power)
        M = sigma_s.shape[0]
        gs = torch.zeros((M, 3), device="cuda")
        ga = torch.zeros((M, 3), device="cuda")
        gpp = torch.zeros((scene.w.photons.count, 3), device="cuda") if scene.bre is not None else None
        native.adjoint(scene.ctx, scene.bre, scene.beams, scene.qv,
                       g_p.contiguous().float() if scene.bre is not None else None,
                       g_b.contiguous().float() if scene.beams is not None else None,
                       gs, ga, gpp, None)
        gpp_out = gpp.to(photon_power.device) if (gpp is not None and photon_power is not None) else None
        return None, gs.to(sigma_s.device), ga.to(sigma_a.device), gpp_out


def estimate(scene: DiffScene, sigma_s, sigma_a, photon_power):
    return VolumetricEstimate.apply(scene, sigma_s, sigma_a, photon_power)


def media_params(media: np.ndarray, device="cuda"):
    return (torch.tensor(media["scattering"], dtype=torch.float32, device=device),
            torch.tensor(media["absorption"], dtype=torch.float32, device=device))
