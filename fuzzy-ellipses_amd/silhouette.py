"""Shape-from-silhouette optimisation: the forward+backward caller of the rasterizer.

Restates the reference notebook's loop (shape_from_silhouette.ipynb: "jax_objective" = BCE of
the alpha clipped to [3e-8, 1 - 3e-8] against a silhouette mask; optimizers.adam; DegradeLR,
utils/sfs_utils.py:62-131) with an added depth term:
    loss = mean_views [ mean_pixels BCE(clip(alpha), mask) + 0.1 * mean_{mask} |depth - target_depth| ]

One step runs entirely on the device through the library's C ABI: the batched forward of all
views, the objective and its image cotangents (`fe_silhouette_loss`), the hand-written backward
(`fe_backward`) and one Adam update per parameter array (`fe_adam_step`).  The parameters are
the means, the upper triangle of the precision factor and the log-weights.  The learning-rate
schedule (`DegradeLR`) is host logic that reads the loss after each step.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import Renderer, SceneTensors, _lib, _stream
from .scenes import Scene

CLIP = 3e-8


class DegradeLR:
    """Learning-rate schedule with early stopping (utils/sfs_utils.py:62-131).

    After `window + p_window` losses it fits a line to the last `window` losses (normalised
    by their mean) and computes the probability that the true slope is below `slope_less`,
    taking the slope estimate as normal with variance 12 s^2 / window^3 (s^2 = residual
    variance with window - 2 degrees of freedom).  When that probability has stayed below
    `p_thresh` for the last `p_window` fits, progress has stalled: the rate drops tenfold and
    the loss history restarts.  `add` returns True to stop: after more than `max_drops`
    drops, or when a drop follows immediately after the previous one (no progress at all).
    The rate is `init_lr / 10**n_drops`.
    """

    def __init__(self, init_lr: float, p_thresh: float = 5e-2, window: int = 10, p_window: int = 5,
                 slope_less: float = 0.0, max_drops: int = 4):
        if not (init_lr > 0 and 0 < p_thresh < 1):
            raise ValueError("DegradeLR: need init_lr > 0 and 0 < p_thresh < 1")
        self.init_lr = float(init_lr)
        self.p_thresh = float(p_thresh)
        self.window = max(3, int(round(window)))
        self.p_window = max(1, int(round(p_window)))
        self.slope_less = float(slope_less)
        self.max_drops = int(max_drops)
        self.n_drops = 0
        self.history: List[float] = []
        self.probs: List[float] = []
        self.count = 0
        self._len_at_last_drop = self.window + 1

    @property
    def lr(self) -> float:
        return self.init_lr / (10.0 ** self.n_drops)

    def step_func(self, _step: int = 0) -> float:
        return self.lr

    def slope_probability(self) -> float:
        """P(slope < slope_less) for the last `window` losses."""
        y = np.asarray(self.history[-self.window:], dtype=np.float64)
        y = y / y.mean()
        n = self.window
        x = np.arange(n, dtype=np.float64)
        xc = x - x.mean()
        slope = float(np.dot(xc, y - y.mean()) / np.dot(xc, xc))
        resid = y - (y.mean() + slope * xc)
        var = float(np.dot(resid, resid)) / (n - 2)
        var_slope = 12.0 * var / n ** 3
        if var_slope <= 0.0:
            # the normal CDF at +-inf, and undefined (never below the threshold) at 0/0
            num = self.slope_less - slope
            return 1.0 if num > 0.0 else (0.0 if num < 0.0 else math.nan)
        return 0.5 * (1.0 + math.erf((self.slope_less - slope) / math.sqrt(2.0 * var_slope)))

    def add(self, loss: float) -> bool:
        self.count += 1
        self.history.append(float(loss))
        n_hist = len(self.history)
        if n_hist < self.window + self.p_window:
            return False
        self.probs.append(self.slope_probability())
        if not all(p < self.p_thresh for p in self.probs[-self.p_window:]):
            return False
        self.n_drops += 1
        if self.n_drops > self.max_drops:
            return True
        shortest = self.window + self.p_window
        if self._len_at_last_drop == shortest and n_hist == shortest:
            return True
        self._len_at_last_drop = n_hist
        self.history = []
        return False


class SilhouetteFit:
    """Fit `init` to silhouettes and depths rendered from `target` over the same views."""

    def __init__(self, init: Scene, target: Scene, device="cuda", tile: int = 16, batch: int = 64,
                 lr: float = 1e-2, mask_threshold: float = 0.5, w_depth: float = 0.1,
                 betas=(0.9, 0.999), eps: float = 1e-8):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.lib = _lib.load()
        self.views = [(0, c, init.beta2, init.beta3) for c in init.cameras]
        self.renderer = Renderer(tile=tile, batch=batch, mode="culled", device=self.device)
        tgt = SceneTensors.from_scene(target, self.device)
        trend = Renderer(tile=tile, batch=batch, mode="culled", device=self.device)
        outs = trend.forward([tgt], [(0, c, target.beta2, target.beta3) for c in target.cameras])
        self.masks = [(a > mask_threshold).float().contiguous() for (_, a) in outs]
        self.tdepths = [d.clone() for (d, _) in outs]
        self.scene = SceneTensors.from_scene(init, self.device)
        self.params = [self.scene.means, self.scene.prec, self.scene.weights_log]
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.lr, self.betas, self.eps, self.w_depth = float(lr), betas, float(eps), float(w_depth)
        self.t = 0
        self.outs = self.renderer.alloc_outputs(self.views)
        self.d_depth = [torch.empty_like(d) for d, _ in self.outs]
        self.d_alpha = [torch.empty_like(a) for _, a in self.outs]
        self.grads = [tuple(torch.zeros_like(p) for p in self.params)]
        self.loss = torch.zeros((), dtype=torch.float32, device=self.device)
        nv = len(self.views)
        self._sil = (_lib.FeSilView * nv)(*[
            _lib.FeSilView(a.data_ptr(), d.data_ptr(), m.data_ptr(), td.data_ptr(), da.data_ptr(), dd.data_ptr(),
                           a.numel())
            for (d, a), m, td, da, dd in zip(self.outs, self.masks, self.tdepths, self.d_alpha, self.d_depth)])
        nb = ctypes.c_size_t(0)
        _lib.check(self.lib.fe_silhouette_workspace_bytes(nv, ctypes.byref(nb)), "fe_silhouette_workspace_bytes")
        self._sil_ws = torch.empty(int(nb.value), dtype=torch.uint8, device=self.device)

    def objective(self) -> torch.Tensor:
        """Loss of the current images into `self.loss` and the cotangents into d_depth/d_alpha."""
        _lib.check(self.lib.fe_silhouette_loss(_stream(self.device), self._sil, len(self.views), self.w_depth, CLIP,
                                               self.loss.data_ptr(), 0, self._sil_ws.data_ptr(),
                                               self._sil_ws.numel()), "fe_silhouette_loss")
        return self.loss

    def loss_and_grads(self):
        """Forward all views, objective, backward -> (loss tensor, (d_means, d_prec, d_w))."""
        self.renderer.forward([self.scene], self.views, save_for_backward=True, outs=self.outs)
        self.objective()
        for g in self.grads[0]:
            g.zero_()
        self.renderer.backward(self.d_depth, self.d_alpha, grads=self.grads)
        return self.loss, self.grads[0]

    def adam(self, lr: Optional[float] = None) -> None:
        self.t += 1
        lr = self.lr if lr is None else float(lr)
        b1, b2 = self.betas
        s = _stream(self.device)
        for p, g, m, v in zip(self.params, self.grads[0], self.m, self.v):
            _lib.check(self.lib.fe_adam_step(s, p.numel(), p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                                             lr, b1, b2, self.eps, self.t), "fe_adam_step")

    def step(self, lr: Optional[float] = None) -> torch.Tensor:
        """One optimisation step; returns the loss of the parameters before the update (a copy)."""
        loss, _ = self.loss_and_grads()
        out = loss.clone()
        self.adam(lr)
        return out

    def run(self, steps: int = 200, schedule: Optional[DegradeLR] = None) -> Dict[str, object]:
        """`steps` steps at the fixed rate, or under `schedule` (which reads each step's loss on
        the host and may stop early)."""
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        hist = []
        stopped = False
        for _ in range(steps):
            hist.append(self.step(None if schedule is None else schedule.lr))
            if schedule is not None and schedule.add(float(hist[-1])):
                stopped = True
                break
        torch.cuda.synchronize(self.device)
        el = time.perf_counter() - t0
        losses = [float(x) for x in hist]
        n = len(losses)
        out = {"steps": n, "seconds": el, "ms_per_step": 1e3 * el / max(n, 1),
               "frames_per_s": n * len(self.views) / el, "loss_first": losses[0] if losses else math.nan,
               "loss_last": losses[-1] if losses else math.nan, "losses": losses}
        if schedule is not None:
            out.update(lr_drops=schedule.n_drops, stopped_early=stopped, final_lr=schedule.lr)
        return out
