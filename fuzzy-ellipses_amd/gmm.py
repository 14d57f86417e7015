"""Gaussian-mixture log-likelihood of a point cloud on the GPU, and a point-cloud fit.

`log_likelihood(params, points)` has the signature and results of the reference's
baseline/fm_render.py:231-259: params = (means (N,3), prec_full (N,3,3), weights_log (N,)),
of which the upper triangle of each precision factor is read; it returns the negative mean
per-point log-likelihood and the (N, M) matrix of per-(Gaussian, point) log terms.  All
arithmetic runs in the library's HIP kernels (fe_gmm_forward / fe_gmm_backward).

`GMMFit` maximises the likelihood of a point cloud with Adam (fe_adam_step): the same
per-(point, Gaussian) quadratic forms as the renderer, used to fit a mixture to 3-D points.
"""
from __future__ import annotations

import ctypes
import time
from typing import Dict, Optional, Tuple

import torch

from . import _lib, _stream


def _f32(x, device) -> torch.Tensor:
    return torch.as_tensor(x, dtype=torch.float32, device=device).contiguous()


class GMM:
    """Device-resident mixture parameters with forward/backward of the point-cloud NLL."""

    def __init__(self, means, prec, weights_log, device="cuda"):
        self.device = torch.device(device)
        self.means = _f32(means, self.device)
        self.prec = _f32(prec, self.device)
        self.weights_log = _f32(weights_log, self.device)
        n = self.means.shape[0]
        if self.means.shape != (n, 3) or self.prec.shape != (n, 3, 3) or self.weights_log.shape != (n,):
            raise ValueError("GMM: expected means (N,3), prec (N,3,3), weights_log (N,)")
        self.lib = _lib.load()
        self._ws: Optional[torch.Tensor] = None
        self._m = -1

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    def _workspace(self, m: int) -> torch.Tensor:
        nb = ctypes.c_size_t(0)
        _lib.check(self.lib.fe_gmm_workspace_bytes(self.n, m, ctypes.byref(nb)), "fe_gmm_workspace_bytes")
        if self._ws is None or self._ws.numel() < nb.value:
            self._ws = torch.empty(int(nb.value), dtype=torch.uint8, device=self.device)
        return self._ws

    def forward(self, points: torch.Tensor, return_res: bool = False, lse: Optional[torch.Tensor] = None,
                nll: Optional[torch.Tensor] = None):
        """(nll scalar tensor, per-point logsumexp (M,), res (N, M) or None)."""
        pts = points if (points.dtype == torch.float32 and points.is_contiguous()) else _f32(points, self.device)
        m = int(pts.shape[0])
        if pts.shape != (m, 3) or pts.device != self.device:
            raise ValueError("GMM.forward: points must be (M,3) on the mixture's device")
        ws = self._workspace(m)
        lse = torch.empty(m, dtype=torch.float32, device=self.device) if lse is None else lse
        nll = torch.empty((), dtype=torch.float32, device=self.device) if nll is None else nll
        res = torch.empty((self.n, m), dtype=torch.float32, device=self.device) if return_res else None
        _lib.check(self.lib.fe_gmm_forward(_stream(self.device), self.n, self.means.data_ptr(), self.prec.data_ptr(),
                                           self.weights_log.data_ptr(), m, pts.data_ptr(), lse.data_ptr(),
                                           None if res is None else res.data_ptr(), nll.data_ptr(), ws.data_ptr(),
                                           ws.numel()), "fe_gmm_forward")
        self._last = (pts, lse)
        return nll, lse, res

    def backward(self, d_nll: float = 1.0, grads=None):
        """Accumulate d_nll * d(nll)/d(means, prec upper triangle, weights_log) of the last forward."""
        pts, lse = self._last
        if grads is None:
            grads = (torch.zeros_like(self.means), torch.zeros_like(self.prec), torch.zeros_like(self.weights_log))
        ws = self._ws
        _lib.check(self.lib.fe_gmm_backward(_stream(self.device), self.n, self.prec.data_ptr(),
                                            self.weights_log.data_ptr(), pts.shape[0], pts.data_ptr(), lse.data_ptr(),
                                            float(d_nll), grads[0].data_ptr(), grads[1].data_ptr(),
                                            grads[2].data_ptr(), ws.data_ptr(), ws.numel()), "fe_gmm_backward")
        return grads


def log_likelihood(params, points, device="cuda") -> Tuple[torch.Tensor, torch.Tensor]:
    """Reference-shaped call: (-mean_j logsumexp_i res[i, j], res (N, M))."""
    means, prec_full, weights_log = params
    g = GMM(means, prec_full, weights_log, device=device)
    nll, _, res = g.forward(_f32(points, g.device), return_res=True)
    return nll, res


class GMMFit:
    """Fit a mixture to a point cloud by Adam on (means, prec upper triangle, weights_log)."""

    def __init__(self, gmm: GMM, points, lr: float = 1e-2, betas=(0.9, 0.999), eps: float = 1e-8):
        self.gmm = gmm
        self.points = _f32(points, gmm.device)
        self.params = [gmm.means, gmm.prec, gmm.weights_log]
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.grads = tuple(torch.zeros_like(p) for p in self.params)
        self.lr, self.betas, self.eps = float(lr), betas, float(eps)
        self.t = 0
        self.lse = torch.empty(self.points.shape[0], dtype=torch.float32, device=gmm.device)
        self.nll = torch.empty((), dtype=torch.float32, device=gmm.device)

    def step(self) -> torch.Tensor:
        g = self.gmm
        g.forward(self.points, lse=self.lse, nll=self.nll)
        out = self.nll.clone()
        for t in self.grads:
            t.zero_()
        g.backward(1.0, self.grads)
        self.t += 1
        s = _stream(g.device)
        for p, gr, m, v in zip(self.params, self.grads, self.m, self.v):
            _lib.check(g.lib.fe_adam_step(s, p.numel(), p.data_ptr(), gr.data_ptr(), m.data_ptr(), v.data_ptr(),
                                          self.lr, self.betas[0], self.betas[1], self.eps, self.t), "fe_adam_step")
        return out

    def run(self, steps: int) -> Dict[str, object]:
        torch.cuda.synchronize(self.gmm.device)
        t0 = time.perf_counter()
        hist = [self.step() for _ in range(steps)]
        torch.cuda.synchronize(self.gmm.device)
        el = time.perf_counter() - t0
        losses = [float(x) for x in hist]
        return {"steps": steps, "seconds": el, "ms_per_step": 1e3 * el / max(steps, 1),
                "nll_first": losses[0], "nll_last": losses[-1], "losses": losses}
