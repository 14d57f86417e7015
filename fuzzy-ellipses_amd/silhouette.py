"""Shape-from-silhouette optimisation: the forward+backward caller of the rasterizer.

Restates the objective of the reference notebook (shape_from_silhouette.ipynb, "jax_objective":
BCE of the clipped alpha against a silhouette mask, clip 3e-8) with an added depth term:
    loss = mean_views [ mean_pixels BCE(clip(alpha, 3e-8, 1-3e-8), mask)
                        + 0.1 * mean_{mask} |depth - target_depth| ]
Gradients of the images come from closed-form derivatives of this loss; gradients of the
Gaussians come from the hand-written backward kernel; parameters (means, upper triangle of
the precision factor, log-weights) are updated with Adam on the device.
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import Renderer, SceneTensors
from .scenes import Scene

CLIP = 3e-8


def loss_and_cotangents(alpha: torch.Tensor, depth: torch.Tensor, mask: torch.Tensor, tdepth: torch.Tensor,
                        w_depth: float = 0.1, n_views: int = 1):
    """Loss of one view (already divided by n_views) and d(loss)/d(depth), d(loss)/d(alpha)."""
    P = alpha.numel()
    a = alpha.clamp(CLIP, 1.0 - CLIP)
    inside = (alpha > CLIP) & (alpha < 1.0 - CLIP)
    bce = -(mask * torch.log(a) + (1.0 - mask) * torch.log1p(-a)).sum() / P
    msum = mask.sum().clamp_min(1.0)
    diff = depth - tdepth
    l1 = (mask * diff.abs()).sum() / msum
    loss = (bce + w_depth * l1) / n_views
    d_alpha = torch.where(inside, -(mask / a - (1.0 - mask) / (1.0 - a)) / P, torch.zeros_like(a)) / n_views
    d_depth = (w_depth * torch.sign(diff) * mask / msum) / n_views
    return loss, d_depth.contiguous(), d_alpha.contiguous()


class Adam:
    def __init__(self, params: Sequence[torch.Tensor], lr: float = 1e-2, betas=(0.9, 0.999), eps: float = 1e-8):
        self.params = list(params)
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.t = 0

    @torch.no_grad()
    def step(self, grads: Sequence[torch.Tensor]) -> None:
        self.t += 1
        c1 = 1.0 - self.b1 ** self.t
        c2 = 1.0 - self.b2 ** self.t
        for p, g, m, v in zip(self.params, grads, self.m, self.v):
            m.mul_(self.b1).add_(g, alpha=1.0 - self.b1)
            v.mul_(self.b2).addcmul_(g, g, value=1.0 - self.b2)
            p.addcdiv_(m / c1, (v / c2).sqrt_().add_(self.eps), value=-self.lr)


class SilhouetteFit:
    """Fit `init` to silhouettes and depths rendered from `target` over the same views."""

    def __init__(self, init: Scene, target: Scene, device="cuda", tile: int = 16, batch: int = 64,
                 lr: float = 1e-2, mask_threshold: float = 0.5):
        self.device = torch.device(device)
        self.views = [(0, c, init.beta2, init.beta3) for c in init.cameras]
        self.renderer = Renderer(tile=tile, batch=batch, mode="culled", device=self.device)
        tgt = SceneTensors.from_scene(target, self.device)
        trend = Renderer(tile=tile, batch=batch, mode="culled", device=self.device)
        outs = trend.forward([tgt], [(0, c, target.beta2, target.beta3) for c in target.cameras])
        self.masks = [(a > mask_threshold).float() for (_, a) in outs]
        self.tdepths = [d.clone() for (d, _) in outs]
        self.scene = SceneTensors.from_scene(init, self.device)
        self.opt = Adam([self.scene.means, self.scene.prec, self.scene.weights_log], lr=lr)
        self.outs = self.renderer.alloc_outputs(self.views)
        self.grads = [(torch.zeros_like(self.scene.means), torch.zeros_like(self.scene.prec),
                       torch.zeros_like(self.scene.weights_log))]

    def loss_and_grads(self):
        """Forward all views, loss, backward -> (loss tensor, (d_means, d_prec, d_w))."""
        r = self.renderer
        r.forward([self.scene], self.views, save_for_backward=True, outs=self.outs)
        n = len(self.views)
        total = torch.zeros((), device=self.device)
        dds, das = [], []
        for (d, a), m, td in zip(self.outs, self.masks, self.tdepths):
            l, dd, da = loss_and_cotangents(a, d, m, td, n_views=n)
            total = total + l
            dds.append(dd)
            das.append(da)
        for g in self.grads[0]:
            g.zero_()
        r.backward(dds, das, grads=self.grads)
        # only the upper triangle of the precision factor is a parameter
        self.grads[0][1].copy_(torch.triu(self.grads[0][1]))
        return total, self.grads[0]

    def step(self) -> torch.Tensor:
        loss, g = self.loss_and_grads()
        self.opt.step(g)
        return loss

    def run(self, steps: int = 200) -> Dict[str, object]:
        losses: List[float] = []
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        hist = []
        for _ in range(steps):
            hist.append(self.step())
        torch.cuda.synchronize(self.device)
        el = time.perf_counter() - t0
        losses = [float(x) for x in hist]
        return {"steps": steps, "seconds": el, "ms_per_step": 1e3 * el / max(steps, 1),
                "frames_per_s": steps * len(self.views) / el, "loss_first": losses[0] if losses else math.nan,
                "loss_last": losses[-1] if losses else math.nan, "losses": losses}
