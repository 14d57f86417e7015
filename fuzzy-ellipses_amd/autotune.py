"""Autotuner for the rasterizer's launch shape: tile edge {8, 16, 32} x LDS batch depth
{32, 64, 128, 256}.

For a workload key (Gaussian-count bucket, resolution, with or without backward) it times
every configuration on the device with HIP events around the real batched forward (and
backward) call, keeps the fastest, and stores it in a JSON tuning table keyed by device name
so later runs reuse it.  Tile size and batch depth change only the evaluation order and the
culling granularity, not the result beyond float32 rounding (checked by the tests).
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import torch

TILES = (8, 16, 32)
BATCHES = (32, 64, 128, 256)
DEFAULT = (16, 64)


def n_bucket(n: int) -> int:
    """Power-of-two bucket of the Gaussian count (16, 32, ..., 8192, ...)."""
    return 1 << max(4, int(math.ceil(math.log2(max(n, 1)))))


def workload_key(scenes, views, backward: bool) -> str:
    """Key of a batch: bucket of its largest scene, its most common resolution, fwd/fwd+bwd."""
    nmax = max(s.n for s in scenes)
    res: Dict[Tuple[int, int], int] = {}
    for (_, cam, _, _) in views:
        res[(cam.H, cam.W)] = res.get((cam.H, cam.W), 0) + cam.H * cam.W
    H, W = max(res, key=res.get)
    return f"n{n_bucket(nmax)}_{H}x{W}_{'fb' if backward else 'f'}"


def _table_path(device) -> str:
    name = torch.cuda.get_device_name(device).replace(" ", "_").replace("/", "_")
    root = os.environ.get("FE_AMD_TUNING_DIR", os.path.join(os.path.expanduser("~"), ".cache", "fuzzy_ellipses_amd"))
    return os.path.join(root, f"tuning_{name}.json")


class Autotuner:
    def __init__(self, device="cuda", path: Optional[str] = None, reps: int = 3):
        self.device = torch.device(device)
        self.reps = reps
        self.path = path
        self.table: Dict[str, dict] = {}
        if self.path is None:
            try:
                self.path = _table_path(self.device)
            except Exception:
                self.path = None
        if self.path and os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    def save(self) -> None:
        if not self.path:
            return
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
        except OSError:
            pass

    def time_config(self, tile: int, batch: int, scenes, views, backward: bool, cot=None) -> float:
        """Median ms of one batched call (forward, plus backward when asked)."""
        from . import Renderer
        r = Renderer(tile=tile, batch=batch, mode="culled", device=self.device)
        outs = r.alloc_outputs(views)
        if backward and cot is None:
            g = torch.Generator(device="cpu").manual_seed(0)
            cot = ([torch.randn(c.H, c.W, generator=g).to(self.device) for (_, c, _, _) in views],
                   [torch.randn(c.H, c.W, generator=g).to(self.device) for (_, c, _, _) in views])
        grads = None

        def call():
            nonlocal grads
            r.forward(scenes, views, save_for_backward=backward, outs=outs)
            if backward:
                if grads is None:
                    grads = [(torch.zeros_like(s.means), torch.zeros_like(s.prec), torch.zeros_like(s.weights_log))
                             for s in scenes]
                r.backward(cot[0], cot[1], grads=grads)

        call()
        stream = torch.cuda.current_stream(self.device)
        times = []
        for _ in range(self.reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            call()
            b.record(stream)
            b.synchronize()
            times.append(a.elapsed_time(b))
        times.sort()
        del r
        return times[len(times) // 2]

    def tune(self, scenes, views, backward: bool = False, tiles: Sequence[int] = TILES,
             batches: Sequence[int] = BATCHES, force: bool = False) -> Tuple[int, int]:
        key = workload_key(scenes, views, backward)
        if not force and key in self.table:
            e = self.table[key]
            return int(e["tile"]), int(e["batch"])
        results: List[Tuple[float, int, int]] = []
        for t in tiles:
            for b in batches:
                results.append((self.time_config(t, b, scenes, views, backward), t, b))
        results.sort()
        best = results[0]
        self.table[key] = {"tile": best[1], "batch": best[2], "ms": best[0],
                           "sweep": {f"{t}x{b}": round(ms, 4) for ms, t, b in results}}
        self.save()
        return best[1], best[2]

    def best(self, scenes, views, backward: bool = False) -> Tuple[int, int]:
        """The tuned (tile, batch) for this workload, or the default if not tuned yet."""
        e = self.table.get(workload_key(scenes, views, backward))
        return (int(e["tile"]), int(e["batch"])) if e else DEFAULT
